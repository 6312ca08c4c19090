// ttc_pme.hip -- smooth PME on the device (replaces pme_qgrid.c,
// pme_calc.c, pme_spline.c) plus the MD loops' kick/drift that feeds it.
//
//   k_md_drift      first half of integrate_vverlet.c:52-83 (kick, drift,
//                   pbc_all.c:23-88) of the MD loops, one thread per site,
//                   with copies of the spread's inputs, so the spread runs
//                   on the side stream beside the step's rebuild.
//   k_spread        charge_grid (pme_qgrid.c:26-113): B-spline weights
//                   (pme_spline.c:10-49) in FP32, q*Mx*My*Mz spread with
//                   int64 fixed-point atomics, so the grid is the same
//                   whatever order the atomics land in.
//   k_fft<MODE>     3-D complex FFT as three line passes in LDS (Stockham,
//                   mixed radix 2/3/4 + direct DFT for other primes).  The
//                   z pass reads the fixed-point grid and clears it for the
//                   next step; the x pass does forward FFT, the influence
//                   function of pme_calc.c:52-93 (energy + virial on demand)
//                   and the inverse FFT without leaving LDS.
//   k_gather        pme_force (pme_calc.c:122-265) + bonds (cbond.c:65-120)
//                   per site; per-block sums of the PME force for the mean
//                   force removal of pme_calc.c:268-277.
#include "ttc_common.h"

#include <math.h>

#include <algorithm>

// FFT lines per workgroup: 8 keep a workgroup's LDS (13 KiB at K = 96)
// small enough to sit beside two pair-kernel groups on a CU
#define FFT_LB 8
#define FFT_THREADS 256
// threads of one spread workgroup (one brick)
#ifndef SPREAD_THREADS
#define SPREAD_THREADS 512
#endif

namespace {

struct Spread {
    double lx, ly, lz;
    int K0, K1, K2;
};

// pme_spline.c:10-49 in FP32: M[1..P] weights, D[1..P] derivatives.
template <int P>
__device__ __forceinline__ void bspline(float w, float *M, float *D) {
    float a[P + 1];
#pragma unroll
    for (int i = 0; i <= P; i++) a[i] = 0.0f;
    a[2] = w;
    a[1] = 1.0f - w;
#pragma unroll
    for (int i = 3; i <= P - 1; i++) {
        float div = 1.0f / (float)(i - 1);
        a[i] = div * w * a[i - 1];
#pragma unroll
        for (int j = 1; j <= i - 2; j++)
            a[i - j] = div * ((w + (float)j) * a[i - j - 1] + ((float)i - (float)j - w) * a[i - j]);
        a[1] *= div * (1.0f - w);
    }
    D[1] = -a[1];
#pragma unroll
    for (int i = 2; i <= P; i++) D[i] = a[i - 1] - a[i];
    float div = 1.0f / (float)(P - 1);
    a[P] = div * w * a[P - 1];
#pragma unroll
    for (int i = 1; i <= P - 2; i++)
        a[P - i] = div * ((w + (float)i) * a[P - i - 1] + ((float)P - (float)i - w) * a[P - i]);
    a[1] *= div * (1.0f - w);
#pragma unroll
    for (int i = 1; i <= P; i++) M[i] = a[i];
}

// a[1..Q] = order-Q cardinal B-spline weights at fractional coordinate w,
// the recursion of pme_spline.c:10-49 run to order Q (Q >= 2)
template <int Q>
__device__ __forceinline__ void bspline_m(float w, float *a) {
#pragma unroll
    for (int i = 0; i <= Q; i++) a[i] = 0.0f;
    a[2] = w;
    a[1] = 1.0f - w;
#pragma unroll
    for (int i = 3; i <= Q; i++) {
        float div = 1.0f / (float)(i - 1);
        a[i] = div * w * a[i - 1];
#pragma unroll
        for (int j = 1; j <= i - 2; j++)
            a[i - j] = div * ((w + (float)j) * a[i - j - 1] + ((float)i - (float)j - w) * a[i - j]);
        a[1] *= div * (1.0f - w);
    }
}

// second derivatives in the same index order: M''_P(x) = M_{P-2}(x)
// - 2 M_{P-2}(x-1) + M_{P-2}(x-2)
template <int P>
__device__ __forceinline__ void bspline_d2(float w, float *D2) {
    float a[P + 1];
    bspline_m<P - 2>(w, a);
#pragma unroll
    for (int k = 1; k <= P; k++) {
        float am2 = (k - 2 >= 1) ? a[k - 2] : 0.0f;
        float am1 = (k - 1 >= 1 && k - 1 <= P - 2) ? a[k - 1] : 0.0f;
        float a0 = (k <= P - 2) ? a[k] : 0.0f;
        D2[k] = am2 - 2.0f * am1 + a0;
    }
}

// grid coordinate u = K (x/L + 0.5) (pme_qgrid.c:26-34); fl = (int)u.
// Wrapped positions give 0 <= u <= K, where floor equals the reference's
// truncation; the base index is reduced modulo K so that a site exactly at
// +L/2 (u = K), or an unwrapped position, still addresses the grid.
__device__ __forceinline__ void grid_coord(double x, double l, int k, int *fl, float *w) {
    double u = (double)k * (x / l + 0.5);
    double fu = floor(u);
    int f = (int)fu;
    f %= k;
    if (f < 0) f += k;
    *fl = f;
    *w = (float)(u - fu);
}

__device__ __forceinline__ int wrapk(int g, int k) {
    if (g < 0) g += k;
    if (g >= k) g -= k;
    return g;
}

template <int P>
__device__ __forceinline__ void spread_site(double4 p, const Spread &sp, long long *qfix) {
    int fx, fy, fz;
    float wx, wy, wz;
    grid_coord(p.x, sp.lx, sp.K0, &fx, &wx);
    grid_coord(p.y, sp.ly, sp.K1, &fy, &wy);
    grid_coord(p.z, sp.lz, sp.K2, &fz, &wz);
    float Mx[P + 1], My[P + 1], Mz[P + 1], D[P + 1];
    bspline<P>(wx, Mx, D);
    bspline<P>(wy, My, D);
    bspline<P>(wz, Mz, D);
    float q = (float)p.w;
#pragma unroll
    for (int xi = 1; xi <= P; xi++) {
        int gx = wrapk(fx - P + xi, sp.K0);
        float qx = q * Mx[xi];
#pragma unroll
        for (int yi = 1; yi <= P; yi++) {
            int gy = wrapk(fy - P + yi, sp.K1);
            float qxy = qx * My[yi];
            size_t rowb = ((size_t)gx * sp.K1 + gy) * sp.K2;
#pragma unroll
            for (int zi = 1; zi <= P; zi++) {
                int gz = wrapk(fz - P + zi, sp.K2);
                long long v = __double2ll_rn((double)(qxy * Mz[zi]) * TTC_QFIX_SCALE);
                atomicAdd((unsigned long long *)&qfix[rowb + gz], (unsigned long long)v);
            }
        }
    }
}

// pbc_all.c:23-88: shift by whole box lengths until inside [-h, h]
__device__ __forceinline__ double pbc1(double x, double l) {
    double h = 0.5 * l;
    for (int it = 0; it < 64 && (x < -h || x > h); it++) {
        double sh = 0.0;
        if (x < -h) sh = +l;
        if (x > h) sh = -l;
        x += sh;
    }
    return x;
}

// Bricks of neighbour cells for the LDS-privatised spread.  A brick's
// tile covers every grid point its sites can touch: the brick's cells
// widened by `delta` (sites drift at most (rclist-rc)/4 plus one step from
// the cell they were sorted into before the trigger rebuilds the list).
// cells per brick edge: at most 2, so a brick has at most 4 (y, z) rows
#define BRICK_ROWS 4
// LDS bytes per site of one spread chunk: q Mx, My, Mz (3P floats) and the
// packed stencil base
#define SPREAD_REC_BYTES(P) (3 * (P) * 4 + 4)

struct Brick {
    int ncx, ncy, ncz;
    int bx, by, bz;
    int nbx, nby, nbz;
    double cwx, cwy, cwz;
    double delta;
    int chunk;     // sites per pass (their weights sit in LDS after the tile)
    int rec_off;   // start of the weights, in 8-byte words after the tile
};

__device__ __forceinline__ void tile_range(int c0, int c1, double cw, double l, int k, double delta, int P,
                                           int *org, int *ext) {
    double xlo = -0.5 * l + c0 * cw - delta, xhi = -0.5 * l + c1 * cw + delta;
    int glo = (int)floor((double)k * (xlo / l + 0.5));
    int ghi = (int)floor((double)k * (xhi / l + 0.5));
    *org = glo - P + 1;
    *ext = ghi - glo + P;
}

// local stencil base of a site inside [org, org+ext), trying the periodic
// images of its grid coordinate; -1 if it lies outside the tile
__device__ __forceinline__ int tile_base(int fl, int org, int ext, int k, int P) {
    int b = fl - P + 1 - org;
    if (b >= 0 && b + P <= ext) return b;
    b += k;
    if (b >= 0 && b + P <= ext) return b;
    b -= 2 * k;
    if (b >= 0 && b + P <= ext) return b;
    return -1;
}

// inputs of the fused second half kick: pair force, gather force, the
// gather's per-block sums (mean PME force), site count, velocity scale
struct KickIn {
    const double4 *fpair = nullptr, *g4 = nullptr;
    const double *partials = nullptr;
    int maxb = 0, nbg = 0, n = 0;
    double vscal = 1.0;
    // original numbering: g4 is read at orig[s]; the new position + charge
    // is written to snap[orig[s]] when snap is set
    const int *orig = nullptr;
    double4 *snap = nullptr;
    long long *fixj = nullptr;   // partner forces of a half-list pair kernel
    int4 *posq = nullptr;        // fixed-point coordinates of the new positions (ttc_posq), when set
    double qkx = 0.0, qky = 0.0, qkz = 0.0;
};

// kick + drift + pbc of one site (integrate_vverlet.c:52-83), its
// original-order snapshot and fixed-point coordinates when asked for, and the
// DLIST displacement check of nblist.c:607-647 on the new position (pcheck =
// {lx, ly, lz, (rclist-rc)/4}; it sets the flag the conditional rebuild that
// follows in the same step reads, so the forces of this step are evaluated
// with a list valid for these positions).
// KICK: the previous step's second half kick (integrate_vverlet.c:93-102)
// first, on the force combined here from the pair and gather outputs (kept
// in f4), exactly as the separate kick kernel computes it.
template <bool KICK>
__device__ __forceinline__ double4 drift_site(int s, double4 *pos4, double4 *vel4, double4 *f4, double dt,
                                              double vscal, const Spread &sp, int *status,
                                              const double4 *__restrict__ anchor, const double *__restrict__ pcheck,
                                              const KickIn &ki, const double *mean) {
    double4 p = pos4[s];
    double4 v = vel4[s];
    double4 f;
    double dth = 0.5 * dt;
    double im = v.w;   // 100/m
    const int so = ki.orig ? ki.orig[s] : s;
    if (KICK) {
        f = ttc_combine(s, 7, ki.fpair, ki.g4, mean, so, ki.fixj, ki.n);
        f4[s] = f;
        v.x = v.x + ki.vscal * dth * v.w * f.x;
        v.y = v.y + ki.vscal * dth * v.w * f.y;
        v.z = v.z + ki.vscal * dth * v.w * f.z;
    } else {
        f = f4[s];
    }
    v.x = vscal * (v.x + dth * im * f.x);
    v.y = vscal * (v.y + dth * im * f.y);
    v.z = vscal * (v.z + dth * im * f.z);
    p.x = pbc1(p.x + dt * v.x, sp.lx);
    p.y = pbc1(p.y + dt * v.y, sp.ly);
    p.z = pbc1(p.z + dt * v.z, sp.lz);
    vel4[s] = v;
    pos4[s] = p;
    if (ki.snap) ki.snap[so] = p;
    if (ki.posq) ki.posq[s] = ttc_posq(p, ki.qkx, ki.qky, ki.qkz);
    double4 a = anchor[s];
    double rcut = pcheck[3];
    double dx = fabs(p.x - a.x);
    if (dx > 0.5 * sp.lx) dx -= sp.lx;
    double dy = fabs(p.y - a.y);
    if (dy > 0.5 * sp.ly) dy -= sp.ly;
    double dz = fabs(p.z - a.z);
    if (dz > 0.5 * sp.lz) dz -= sp.lz;
    if (dx > rcut || dy > rcut || dz > rcut || ttc_r2_exact(dx, dy, dz) > rcut * rcut)
        atomicOr(&status[ST_NLFLAG], 1);
    return p;
}

// charge_grid (pme_qgrid.c:77-113) of every site of a brick into an LDS
// tile of int64 fixed point, flushed to the global fixed-point grid with
// one atomic per touched grid point.  Integer adds make the grid
// independent of order.
template <int P>
__device__ __forceinline__ void spread_body(int b, const double4 *__restrict__ pos4, const Spread &sp,
                                            const Brick &br, const int *__restrict__ cell_start,
                                            const int *__restrict__ cell_cnt, long long *qfix) {
    extern __shared__ unsigned long long tile[];
    // one chunk of sites after the tile: q Mx[1..P], My[1..P], Mz[1..P] per
    // site and the packed tile-local stencil base (x | y << 10 | z << 20;
    // -1: site spread straight to the grid)
    const int chunk = br.chunk;
    float *rw = (float *)(tile + br.rec_off);
    int *rb = (int *)(rw + chunk * 3 * P);
    int ibx = b % br.nbx, iby = (b / br.nbx) % br.nby, ibz = b / (br.nbx * br.nby);
    int cx0 = ibx * br.bx, cy0 = iby * br.by, cz0 = ibz * br.bz;
    int cx1 = min(cx0 + br.bx, br.ncx), cy1 = min(cy0 + br.by, br.ncy), cz1 = min(cz0 + br.bz, br.ncz);
    int ox, oy, oz, tx, ty, tz;
    tile_range(cx0, cx1, br.cwx, sp.lx, sp.K0, br.delta, P, &ox, &tx);
    tile_range(cy0, cy1, br.cwy, sp.ly, sp.K1, br.delta, P, &oy, &ty);
    tile_range(cz0, cz1, br.cwz, sp.lz, sp.K2, br.delta, P, &oz, &tz);
    int tsz = tx * ty * tz;
    for (int t = threadIdx.x; t < tsz; t += blockDim.x) tile[t] = 0ull;
    // the brick's sites are the contiguous site ranges of its (cy, cz) cell
    // rows (cells are sorted x fastest); up to BRICK_ROWS rows, taken as one
    // flat index range
    const int nyr = cy1 - cy0, nrow = nyr * (cz1 - cz0);
    int rs0[BRICK_ROWS], rend[BRICK_ROWS];   // row start, cumulative end
    int tot = 0;
#pragma unroll
    for (int r = 0; r < BRICK_ROWS; r++) {
        rs0[r] = 0;
        if (r < nrow) {
            int cy = cy0 + r % nyr, cz = cz0 + r / nyr;
            int c0 = cx0 + br.ncx * (cy + br.ncy * cz);
            int cl = (cx1 - 1) + br.ncx * (cy + br.ncy * cz);
            rs0[r] = cell_start[c0];
            tot += cell_start[cl] + cell_cnt[cl] - rs0[r];
        }
        rend[r] = tot;
    }
    __syncthreads();
    for (int c0 = 0; c0 < tot; c0 += chunk) {
        // pass 1: one thread per site (B-spline weights)
        for (int i = c0 + threadIdx.x; i < min(tot, c0 + chunk); i += blockDim.x) {
            int s = 0;
#pragma unroll
            for (int r = BRICK_ROWS - 1; r >= 0; r--)
                if (i < rend[r]) s = rs0[r] + i - (r ? rend[r - 1] : 0);
            const double4 p = pos4[s];
            double ux = (double)sp.K0 * (p.x / sp.lx + 0.5), uy = (double)sp.K1 * (p.y / sp.ly + 0.5),
                   uz = (double)sp.K2 * (p.z / sp.lz + 0.5);
            double fxd = floor(ux), fyd = floor(uy), fzd = floor(uz);
            int bxl = tile_base((int)fxd, ox, tx, sp.K0, P);
            int byl = tile_base((int)fyd, oy, ty, sp.K1, P);
            int bzl = tile_base((int)fzd, oz, tz, sp.K2, P);
            const int k = i - c0;
            float *w = rw + k * 3 * P;
            if (bxl < 0 || byl < 0 || bzl < 0) {
                spread_site<P>(p, sp, qfix);   // outside the tile: straight to the grid
                rb[k] = -1;
                continue;
            }
            float Mx[P + 1], My[P + 1], Mz[P + 1], D[P + 1];
            bspline<P>((float)(ux - fxd), Mx, D);
            bspline<P>((float)(uy - fyd), My, D);
            bspline<P>((float)(uz - fzd), Mz, D);
            float q = (float)p.w;
#pragma unroll
            for (int xi = 0; xi < P; xi++) {
                w[xi] = q * Mx[xi + 1];
                w[P + xi] = My[xi + 1];
                w[2 * P + xi] = Mz[xi + 1];
            }
            rb[k] = bxl | (byl << 10) | (bzl << 20);
        }
        __syncthreads();
        // pass 2: P*P threads per site, each one (x, y) row of P z points
        const int nk = min(tot - c0, chunk);
        for (int it = threadIdx.x; it < nk * P * P; it += blockDim.x) {
            const int k = it / (P * P), xy = it - k * (P * P), xi = xy / P, yi = xy - xi * P;
            const int base = rb[k];
            if (base < 0) continue;
            const int bxl = base & 1023, byl = (base >> 10) & 1023, bzl = base >> 20;
            const float *w = rw + k * 3 * P;
            const float qxy = w[xi] * w[P + yi];
            const int rowb = ((bxl + xi) * ty + (byl + yi)) * tz + bzl;
#pragma unroll
            for (int zi = 0; zi < P; zi++) {
                long long v = __double2ll_rn((double)(qxy * w[2 * P + zi]) * TTC_QFIX_SCALE);
                atomicAdd(&tile[rowb + zi], (unsigned long long)v);
            }
        }
        __syncthreads();
    }
    // flush: one global atomic per touched tile point; t -> (lx, ly, lz) by
    // multiply-high with ceil(2^32 / d) (exact: t * d < 2^32)
    const unsigned mz = 0xFFFFFFFFu / (unsigned)tz + 1u, my = 0xFFFFFFFFu / (unsigned)ty + 1u;
    for (int t = threadIdx.x; t < tsz; t += blockDim.x) {
        unsigned long long v = tile[t];
        if (!v) continue;
        int r = (int)__umulhi((unsigned)t, mz), lz = t - r * tz;
        int lx = (int)__umulhi((unsigned)r, my), ly = r - lx * ty;
        int gx = (ox + lx) % sp.K0, gy = (oy + ly) % sp.K1, gz = (oz + lz) % sp.K2;
        if (gx < 0) gx += sp.K0;
        if (gy < 0) gy += sp.K1;
        if (gz < 0) gz += sp.K2;
        atomicAdd((unsigned long long *)&qfix[((size_t)gx * sp.K1 + gy) * sp.K2 + gz], v);
    }
}

struct SpreadPack {
    const double4 *pos4;
    Spread sp;
    Brick br;
    const int *cell_start, *cell_cnt;
    long long *qfix;
};

template <int P>
__global__ void __launch_bounds__(SPREAD_THREADS) k_spread(SpreadPack a) {
    spread_body<P>(blockIdx.x, a.pos4, a.sp, a.br, a.cell_start, a.cell_cnt, a.qfix);
}
template <int P>
__global__ void __launch_bounds__(SPREAD_THREADS) kb_spread(const SpreadPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const SpreadPack a = pk[m.rep];
    spread_body<P>(m.blk, a.pos4, a.sp, a.br, a.cell_start, a.cell_cnt, a.qfix);
}

// An MD loop's drift alone, one thread per site of the sorted order
// (drift_site; with the fused kick in the velocity-Verlet loop), which also
// copies what this step's spread reads: the new position + charge (spos)
// and the cell ranges of the current order (scs, scc).  The spread then
// runs on the side stream beside the step's conditional rebuild, which
// reorders pos4 and rewrites the cell ranges, and the pair kernel waits for
// the drift only.
struct MdDriftPack {
    double4 *pos4, *vel4, *f4;
    double dt, vscal;
    Spread sp;
    int n, ncell;
    const int *cell_start, *cell_cnt;
    int *scs, *scc;
    double4 *spos;
    int *status;
    const double4 *anchor;
    const double *pcheck;
    KickIn ki;
};

template <bool KICK>
__device__ __forceinline__ void md_drift_body(int blk, const MdDriftPack &a) {
    __shared__ double mean[3];
    // a velocity-Verlet loop whose rebuild overflowed stops here; the host
    // grows the list and redoes that step from its drifted positions (the
    // NHC loops, without a snapshot, redo no step)
    if (a.ki.snap && (a.status[ST_OVERFLOW] & TTC_OVF_REDO)) return;
    if (KICK) ttc_mean_pme(a.ki.partials, a.ki.maxb, a.ki.nbg, a.ki.n, mean);
    const int t = blk * blockDim.x + threadIdx.x;
    if (t < a.ncell) {
        a.scs[t] = a.cell_start[t];
        a.scc[t] = a.cell_cnt[t];
    }
    if (t >= a.n) return;
    a.spos[t] = drift_site<KICK>(t, a.pos4, a.vel4, a.f4, a.dt, a.vscal, a.sp, a.status, a.anchor, a.pcheck, a.ki,
                                 mean);
}

template <bool KICK>
__global__ void __launch_bounds__(256) k_md_drift(MdDriftPack a) { md_drift_body<KICK>(blockIdx.x, a); }

// batched (ttc_batch.hip): the NHC step's drift, block `blk` of context `rep`
__global__ void __launch_bounds__(256) kb_md_drift(const MdDriftPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    md_drift_body<false>(m.blk, pk[m.rep]);
}

struct FftArgs {
    int K0, K1, K2;
    int n;          // line length
    int nrad;
    int rad[16];
    int ns[16];             // stage st: product of the radices before it
    int nr[16];             // n / rad[st]
    unsigned mag[16];       // ceil(2^32 / ns[st]) mod 2^32: j / ns[st] by multiply-high
    unsigned mag_n;         // ceil(2^32 / n)
    int H;                  // kz values kept: K2 / 2 + 1 (the grid is real)
    unsigned mag_h;         // ceil(2^32 / H)
    const float2 *tw;
};

// q = j / d for j >= 0, d >= 1 with j * d < 2^32, where m = ceil(2^32 / d)
// truncated to 32 bits (so m = 0 stands for d = 1): umulhi(j, m) =
// floor(j/d + j e / (d 2^32)) with e = m d - 2^32 < d, and j e < 2^32 keeps
// the second term below 1/d, so the floor is unchanged
__device__ __forceinline__ int udiv_mag(int j, unsigned m) { return m ? (int)__umulhi((unsigned)j, m) : j; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// e^{sign 2 pi i m / n}, 0 <= m < n
__device__ __forceinline__ float2 twid(const float2 *tw, int m, int sign) {
    float2 t = tw[m];
    return sign > 0 ? t : make_float2(t.x, -t.y);
}

// Stockham passes over FFT_LB lines held in LDS; element e of line l sits at
// buf[l*ls + e*es].  Work item t takes line t % FFT_LB, so consecutive lanes
// touch consecutive lines: with es = FFT_LB (ls = 1) they read consecutive
// words, and with es = 1 the line stride ls = n + 1 spreads them over the
// banks.  Stage constants come from the host (no integer division here).
// Returns the buffer that holds the result.
__device__ float2 *fft_lds(float2 *A, float2 *B, int ls, int es, const FftArgs &f, const float2 *tw, int sign) {
    const int n = f.n;
    float2 *src = A, *dst = B;
    for (int st = 0; st < f.nrad; st++) {
        const int R = f.rad[st], Ns = f.ns[st], nR = f.nr[st];
        const unsigned mNs = f.mag[st];
        const int tstride = nR / Ns;   // n / (Ns R)
        for (int t = threadIdx.x; t < FFT_LB * nR; t += blockDim.x) {
            const int line = t & (FFT_LB - 1), j = t / FFT_LB;
            const int q = udiv_mag(j, mNs), k = j - q * Ns;
            // radix twiddle indices r * k * tstride stay below n for r < R
            const int tstep = k * tstride;
            const float2 *s0 = src + line * ls;
            float2 *d0 = dst + line * ls;
            const int idxD = q * Ns * R + k;
            if (R == 2) {
                float2 a = s0[j * es], b = cmul(s0[(j + nR) * es], twid(tw, tstep, sign));
                d0[idxD * es] = make_float2(a.x + b.x, a.y + b.y);
                d0[(idxD + Ns) * es] = make_float2(a.x - b.x, a.y - b.y);
            } else if (R == 4) {
                float2 a = s0[j * es];
                float2 b = cmul(s0[(j + nR) * es], twid(tw, tstep, sign));
                float2 c = cmul(s0[(j + 2 * nR) * es], twid(tw, 2 * tstep, sign));
                float2 d = cmul(s0[(j + 3 * nR) * es], twid(tw, 3 * tstep, sign));
                float2 apc = make_float2(a.x + c.x, a.y + c.y), amc = make_float2(a.x - c.x, a.y - c.y);
                float2 bpd = make_float2(b.x + d.x, b.y + d.y), bmd = make_float2(b.x - d.x, b.y - d.y);
                // (sign i) * bmd
                float2 ib = sign > 0 ? make_float2(-bmd.y, bmd.x) : make_float2(bmd.y, -bmd.x);
                d0[idxD * es] = make_float2(apc.x + bpd.x, apc.y + bpd.y);
                d0[(idxD + Ns) * es] = make_float2(amc.x + ib.x, amc.y + ib.y);
                d0[(idxD + 2 * Ns) * es] = make_float2(apc.x - bpd.x, apc.y - bpd.y);
                d0[(idxD + 3 * Ns) * es] = make_float2(amc.x - ib.x, amc.y - ib.y);
            } else if (R == 3) {
                float2 a = s0[j * es];
                float2 b = cmul(s0[(j + nR) * es], twid(tw, tstep, sign));
                float2 c = cmul(s0[(j + 2 * nR) * es], twid(tw, 2 * tstep, sign));
                const float h = 0.86602540378443864676f * (float)sign;
                float2 bpc = make_float2(b.x + c.x, b.y + c.y), bmc = make_float2(b.x - c.x, b.y - c.y);
                float2 m = make_float2(a.x - 0.5f * bpc.x, a.y - 0.5f * bpc.y);
                d0[idxD * es] = make_float2(a.x + bpc.x, a.y + bpc.y);
                d0[(idxD + Ns) * es] = make_float2(m.x - h * bmc.y, m.y + h * bmc.x);
                d0[(idxD + 2 * Ns) * es] = make_float2(m.x + h * bmc.y, m.y - h * bmc.x);
            } else {
                // direct DFT of length R: X_q = sum_r src_r w^{r (tstep + q n/R)}
                for (int qq = 0; qq < R; qq++) {
                    float2 acc = make_float2(0.0f, 0.0f);
                    int base = tstep + qq * nR;
                    for (int r = 0; r < R; r++) {
                        float2 v = cmul(s0[(j + r * nR) * es], twid(tw, (r * base) % n, sign));
                        acc.x += v.x;
                        acc.y += v.y;
                    }
                    d0[(idxD + qq * Ns) * es] = acc;
                }
            }
        }
        __syncthreads();
        float2 *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// The charge grid is real, so its transform is Hermitian, S(-k) = conj S(k),
// and only kz = 0 .. K2/2 (H values) is kept: grid qc[K0][K1][H].
// MODE 0: z lines, forward: two real lines of the fixed-point grid (cleared
//         after reading) as the real and imaginary parts of one complex
//         line, whose transform Z gives both half spectra:
//         A(k) = (Z(k) + conj Z(n-k)) / 2, B(k) = -i (Z(k) - conj Z(n-k)) / 2
// MODE 1: y lines of the half grid, forward (sign +1) / inverse (sign -1)
// MODE 2: x lines of the half grid, forward + influence function + inverse
// MODE 3: z lines, inverse: two half spectra completed by Hermitian symmetry
//         as Z = A + i B, whose inverse holds the two real lines in its real
//         and imaginary parts
// The y and x passes take FFT_LB lines per workgroup from the flat list of
// (fixed coordinate, kz) lines.
template <int MODE, bool ENERGY>
__device__ __forceinline__ void fft_body(int bid, const FftArgs &f, int sign, long long *qfix, float2 *qc, float *qr,
                                         const float *eterm, double fac, double lx, double ly, double lz,
                                         double *partials, int maxb) {
    extern __shared__ float2 lds[];
    const int n = f.n, H = f.H;
    // buffers of FFT_LB lines; z lines (MODE 0, 3) are padded to n + 1
    const int lsz = (MODE == 0 || MODE == 3) ? n + 1 : n;
    float2 *A = lds, *B = lds + FFT_LB * lsz;
    float2 *tw = B + FFT_LB * lsz;
    for (int m = threadIdx.x; m < n; m += blockDim.x) tw[m] = f.tw[m];
    int ls, es;
    size_t r0 = 0;   // first real z line of the workgroup (MODE 0, 3)
    int nreal = 0;   // its real z lines, at most 2 FFT_LB
    // MODE 1, 2: this thread's line (the workgroup's line threadIdx.x % FFT_LB)
    int fix = 0, kz = 0;
    bool lok = false;
    if (MODE == 0 || MODE == 3) {
        r0 = (size_t)bid * 2 * FFT_LB;
        nreal = min(2 * FFT_LB, f.K0 * f.K1 - (int)r0);
        ls = n + 1; es = 1;
        for (int idx = threadIdx.x; idx < FFT_LB * n; idx += blockDim.x) {
            const int l = udiv_mag(idx, f.mag_n), e = idx - l * n;
            const int ra = 2 * l, rb = 2 * l + 1;
            float2 v = make_float2(0.0f, 0.0f);
            if (MODE == 0) {
                if (ra < nreal) {
                    const size_t g = (r0 + ra) * n + e;
                    const long long q = qfix[g];
                    qfix[g] = 0;
                    v.x = (float)((double)q * TTC_QFIX_INV);
                }
                if (rb < nreal) {
                    const size_t g = (r0 + rb) * n + e;
                    const long long q = qfix[g];
                    qfix[g] = 0;
                    v.y = (float)((double)q * TTC_QFIX_INV);
                }
            } else {
                // element e >= H of a half spectrum is conj of element n - e
                const bool lo = e < H;
                const int k = lo ? e : n - e;
                float2 a = ra < nreal ? qc[(r0 + ra) * H + k] : make_float2(0.0f, 0.0f);
                float2 b = rb < nreal ? qc[(r0 + rb) * H + k] : make_float2(0.0f, 0.0f);
                if (!lo) { a.y = -a.y; b.y = -b.y; }
                v = make_float2(a.x - b.y, a.y + b.x);
            }
            A[l * ls + e] = v;
        }
    } else {
        const int nfix = MODE == 1 ? f.K0 : f.K1;
        const int fl = bid * FFT_LB + (int)(threadIdx.x & (FFT_LB - 1));
        lok = fl < nfix * H;
        if (lok) {
            fix = udiv_mag(fl, f.mag_h);
            kz = fl - fix * H;
        }
        ls = 1; es = FFT_LB;
        // lines along y (MODE 1: fixed x) or x (MODE 2: fixed y); blockDim is
        // a multiple of FFT_LB, so idx % FFT_LB is this thread's line
        for (int idx = threadIdx.x; idx < n * FFT_LB; idx += blockDim.x) {
            const int e = idx / FFT_LB;
            float2 v = make_float2(0.0f, 0.0f);
            if (lok) {
                const size_t g = MODE == 1 ? ((size_t)fix * f.K1 + e) * H + kz : ((size_t)e * f.K1 + fix) * H + kz;
                v = qc[g];
            }
            A[idx] = v;
        }
    }
    __syncthreads();
    float2 *res = fft_lds(A, B, ls, es, f, tw, MODE == 3 ? -1 : (MODE == 1 ? sign : +1));

    if (MODE == 2) {
        // influence function (pme_calc.c:52-93); element e = k1, line -> (k2 = fix, k3 = kz).
        // Energy and virial: kz and -kz give equal terms (every term is even
        // in m), so the kept 0 < kz < K2/2 count twice
        double en = 0.0, w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
        const double wgt = (kz == 0 || 2 * kz == f.K2) ? 1.0 : 2.0;
        for (int idx = threadIdx.x; idx < n * FFT_LB; idx += blockDim.x) {
            if (!lok) continue;
            const int e = idx / FFT_LB;
            const size_t g = ((size_t)e * f.K1 + fix) * H + kz;
            float et = eterm[g];
            float2 v = res[idx];
            if (ENERGY && g != 0) {
                int k1 = e, k2 = fix, k3 = kz;
                int mx = k1 - (k1 > f.K0 / 2 ? f.K0 : 0);
                int my = k2 - (k2 > f.K1 / 2 ? f.K1 : 0);
                int mz = k3;
                double hx = mx / lx, hy = my / ly, hz = mz / lz;
                double m2 = hx * hx + hy * hy + hz * hz;
                double st2 = (double)v.x * v.x + (double)v.y * v.y;
                double es2 = wgt * (double)et * st2;
                double wt = -2.0 * (fac * m2 + 1.0) / m2;
                en += es2;
                w0 += es2 * (wt * hx * hx + 1.0);
                w1 += es2 * (wt * hy * hx);
                w2 += es2 * (wt * hy * hy + 1.0);
                w3 += es2 * (wt * hz * hx);
                w4 += es2 * (wt * hz * hy);
                w5 += es2 * (wt * hz * hz + 1.0);
            }
            res[idx] = make_float2(v.x * et, v.y * et);
        }
        if (ENERGY) {
            double vals[7] = {en, w0, w1, w2, w3, w4, w5};
            __shared__ double red[FFT_THREADS / 64][7];
            int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
            for (int q = 0; q < 7; q++) {
                double t = ttc_wave_sum(vals[q]);
                if (lane == 0) red[wid][q] = t;
            }
            __syncthreads();
            if (threadIdx.x < 7) {
                double t = 0.0;
                for (int w = 0; w < FFT_THREADS / 64; w++) t += red[w][threadIdx.x];
                partials[(size_t)(EV_REC + threadIdx.x) * maxb + bid] = t;
            }
        }
        __syncthreads();
        float2 *other = res == A ? B : A;
        res = fft_lds(res, other, ls, es, f, tw, -1);
    }

    // store
    if (MODE == 0) {
        // both half spectra of complex line l
        for (int idx = threadIdx.x; idx < FFT_LB * H; idx += blockDim.x) {
            const int l = udiv_mag(idx, f.mag_h), k = idx - l * H;
            const int ra = 2 * l, rb = 2 * l + 1;
            const float2 zk = res[l * ls + k], zm = res[l * ls + (k ? n - k : 0)];
            if (ra < nreal) qc[(r0 + ra) * H + k] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            if (rb < nreal) qc[(r0 + rb) * H + k] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        }
    } else if (MODE == 3) {
        for (int idx = threadIdx.x; idx < FFT_LB * n; idx += blockDim.x) {
            const int l = udiv_mag(idx, f.mag_n), e = idx - l * n;
            const int ra = 2 * l, rb = 2 * l + 1;
            const float2 z = res[l * ls + e];
            if (ra < nreal) qr[(r0 + ra) * n + e] = z.x;
            if (rb < nreal) qr[(r0 + rb) * n + e] = z.y;
        }
    } else {
        for (int idx = threadIdx.x; idx < n * FFT_LB; idx += blockDim.x) {
            if (!lok) continue;
            const int e = idx / FFT_LB;
            const size_t g = MODE == 1 ? ((size_t)fix * f.K1 + e) * H + kz : ((size_t)e * f.K1 + fix) * H + kz;
            qc[g] = res[idx];
        }
    }
}

struct FftPack {
    FftArgs f;
    int sign;
    long long *qfix;
    float2 *qc;
    float *qr;
    const float *eterm;
    double fac, lx, ly, lz;
    double *partials;
    int maxb;
};

template <int MODE, bool ENERGY>
__global__ void __launch_bounds__(FFT_THREADS) k_fft(FftPack a) {
    fft_body<MODE, ENERGY>(blockIdx.x, a.f, a.sign, a.qfix, a.qc, a.qr, a.eterm, a.fac, a.lx, a.ly, a.lz, a.partials,
                           a.maxb);
}
template <int MODE, bool ENERGY>
__global__ void __launch_bounds__(FFT_THREADS) kb_fft(const FftPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    // the replica's arguments are read in place (a local copy, indexed by
    // stage at run time, would live in scratch memory)
    const FftPack &a = pk[m.rep];
    fft_body<MODE, ENERGY>(m.blk, a.f, a.sign, a.qfix, a.qc, a.qr, a.eterm, a.fac, a.lx, a.ly, a.lz, a.partials,
                           a.maxb);
}

struct GatherConst {
    double lx, ly, lz, hx, hy, hz;
    int K0, K1, K2;
    double qscale;   // QFACTOR / epsRF
};

template <int P, bool ENERGY, bool HES = false>
__device__ __forceinline__ void gather_body(int bid, int n, const double4 *__restrict__ pos4,
                                            const int *__restrict__ orig, const int *__restrict__ inv,
                                            const float *__restrict__ qr, const GatherConst &gc,
                                            const int *__restrict__ b_off, const int *__restrict__ b_part,
                                            const double *__restrict__ b_r0, const double *__restrict__ b_k,
                                            double4 *__restrict__ g4, double *partials, int maxb, int do_pme,
                                            int do_bonds, const float *__restrict__ theta_near,
                                            const int *__restrict__ a_off = nullptr,
                                            const int *__restrict__ a_ref = nullptr,
                                            const double4 *__restrict__ a_tf = nullptr, int amask = 0) {
    constexpr int ND = 2 * P - 1;   // grid offsets -(P-1) .. P-1
    __shared__ float sth[HES ? ND * ND * ND : 1];
    if (HES) {
        for (int i = threadIdx.x; i < ND * ND * ND; i += blockDim.x) sth[i] = theta_near[i];
        __syncthreads();
    }
    int s = bid * blockDim.x + threadIdx.x;
    double px = 0.0, py = 0.0, pz = 0.0;
    double eb = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
    double kh[3] = {0.0, 0.0, 0.0}, bh[3] = {0.0, 0.0, 0.0};
    if (s < n) {
        double4 p = pos4[s];
        if (do_pme) {
        int fx, fy, fz;
        float wx, wy, wz;
        grid_coord(p.x, gc.lx, gc.K0, &fx, &wx);
        grid_coord(p.y, gc.ly, gc.K1, &fy, &wy);
        grid_coord(p.z, gc.lz, gc.K2, &fz, &wz);
        float Mx[P + 1], My[P + 1], Mz[P + 1], Dx[P + 1], Dy[P + 1], Dz[P + 1];
        bspline<P>(wx, Mx, Dx);
        bspline<P>(wy, My, Dy);
        bspline<P>(wz, Mz, Dz);
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
        for (int xi = 1; xi <= P; xi++) {
            int gx = wrapk(fx - P + xi, gc.K0);
#pragma unroll
            for (int yi = 1; yi <= P; yi++) {
                int gy = wrapk(fy - P + yi, gc.K1);
                const float *rowp = qr + ((size_t)gx * gc.K1 + gy) * gc.K2;
                float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
                for (int zi = 1; zi <= P; zi++) {
                    int gz = wrapk(fz - P + zi, gc.K2);
                    float t = rowp[gz];
                    sx += Mz[zi] * t;
                    sz += Dz[zi] * t;
                }
                ax -= Dx[xi] * My[yi] * sx;
                ay -= Mx[xi] * Dy[yi] * sx;
                az -= Mx[xi] * My[yi] * sz;
                (void)sy;
            }
        }
        double f = p.w * gc.qscale;
        px = (double)ax * (f * gc.K0 / gc.lx);
        py = (double)ay * (f * gc.K1 / gc.ly);
        pz = (double)az * (f * gc.K2 / gc.lz);
        if (HES) {
            // d2E/dd_i2 = C q_i (K_d/L_d)^2 [sum_m W''(m) (theta*Q)(m)
            //             + q_i sum_{m,m'} W'(m) theta(m-m') W'(m')]
            float D2x[P + 1], D2y[P + 1], D2z[P + 1];
            bspline_d2<P>(wx, D2x);
            bspline_d2<P>(wy, D2y);
            bspline_d2<P>(wz, D2z);
            float hA[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int xi = 1; xi <= P; xi++) {
                int gx = wrapk(fx - P + xi, gc.K0);
#pragma unroll
                for (int yi = 1; yi <= P; yi++) {
                    int gy = wrapk(fy - P + yi, gc.K1);
                    const float *rowp = qr + ((size_t)gx * gc.K1 + gy) * gc.K2;
                    float sm = 0.0f, sd2 = 0.0f;
#pragma unroll
                    for (int zi = 1; zi <= P; zi++) {
                        float t = rowp[wrapk(fz - P + zi, gc.K2)];
                        sm += Mz[zi] * t;
                        sd2 += D2z[zi] * t;
                    }
                    hA[0] += D2x[xi] * My[yi] * sm;
                    hA[1] += Mx[xi] * D2y[yi] * sm;
                    hA[2] += Mx[xi] * My[yi] * sd2;
                }
            }
            // autocorrelations of the weights and of their derivatives over
            // the stencil, then the sum over grid offsets of theta
            float cm[3][ND], cd[3][ND];
            const float *Mw[3] = {Mx, My, Mz}, *Dw[3] = {Dx, Dy, Dz};
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int o = 0; o < ND; o++) {
                    int delta = o - (P - 1);
                    float am = 0.0f, ad = 0.0f;
#pragma unroll
                    for (int a = 1; a <= P; a++) {
                        int b = a - delta;
                        if (b >= 1 && b <= P) { am += Mw[d][a] * Mw[d][b]; ad += Dw[d][a] * Dw[d][b]; }
                    }
                    cm[d][o] = am;
                    cd[d][o] = ad;
                }
            float hB[3] = {0.0f, 0.0f, 0.0f};
            for (int ox = 0; ox < ND; ox++)
                for (int oy = 0; oy < ND; oy++) {
                    const float *th = sth + (ox * ND + oy) * ND;
                    float tm = 0.0f, td = 0.0f;
#pragma unroll
                    for (int oz = 0; oz < ND; oz++) {
                        tm += th[oz] * cm[2][oz];
                        td += th[oz] * cd[2][oz];
                    }
                    hB[0] += cd[0][ox] * cm[1][oy] * tm;
                    hB[1] += cm[0][ox] * cd[1][oy] * tm;
                    hB[2] += cm[0][ox] * cm[1][oy] * td;
                }
            double kl[3] = {gc.K0 / gc.lx, gc.K1 / gc.ly, gc.K2 / gc.lz};
#pragma unroll
            for (int d = 0; d < 3; d++)
                kh[d] = f * kl[d] * kl[d] * ((double)hA[d] + p.w * (double)hB[d]);
        }
        }
        // bonds of this site (cbond.c:65-120), each bond seen from both ends
        double bx = 0.0, by = 0.0, bz = 0.0;
        int o = orig ? orig[s] : s;   // no map: the sites are in the original numbering
        int kb0 = do_bonds ? b_off[o] : 0, kb1 = do_bonds ? b_off[o + 1] : 0;
        for (int k = kb0; k < kb1; k++) {
            double4 q = pos4[inv ? inv[b_part[k]] : b_part[k]];
            double dx = ttc_minimg(q.x - p.x, gc.hx, gc.lx);
            double dy = ttc_minimg(q.y - p.y, gc.hy, gc.ly);
            double dz = ttc_minimg(q.z - p.z, gc.hz, gc.lz);
            double dr = sqrt(dx * dx + dy * dy + dz * dz);
            double rr = dr - b_r0[k];
            double fr = -b_k[k] * rr * 2.0;
            double fxb = fr / dr * dx, fyb = fr / dr * dy, fzb = fr / dr * dz;
            bx -= fxb; by -= fyb; bz -= fzb;
            if (HES) {
                // cbond2.c:105-122, one end of the bond (the other end adds
                // the same amount: the reference's factor 2)
                double de = -fr / dr;
                double term = (2.0 * b_k[k] - de) / (dr * dr);
                bh[0] += term * dx * dx + de;
                bh[1] += term * dy * dy + de;
                bh[2] += term * dz * dz + de;
            }
            if (ENERGY) {
                eb += -(fr * rr * 0.5) * 0.5;
                v0 += 0.5 * fxb * dx; v1 += 0.5 * fyb * dx; v2 += 0.5 * fyb * dy;
                v3 += 0.5 * fzb * dx; v4 += 0.5 * fzb * dy; v5 += 0.5 * fzb * dz;
            }
        }
        // angle-type terms of this site (their kernel ran just before):
        // the term forces on this site, in the fixed CSR order
        if (amask) {
            for (int k = a_off[o]; k < a_off[o + 1]; k++) {
                unsigned r = (unsigned)a_ref[k];   // slot | kind << 30 (kinds 2, 3 set the sign bit)
                if ((amask >> (r >> 30)) & 1) {
                    double4 f = a_tf[r & 0x3FFFFFFFu];
                    bx += f.x; by += f.y; bz += f.z;
                }
            }
        }
        g4[s] = make_double4(px + bx, py + by, pz + bz, 0.0);
    }
    // per-block sums (PME force sum always; bond energy/virial on demand)
    __shared__ double red[4][16];
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double vals[16] = {px, py, pz, eb, v0, v1, v2, v3, v4, v5, kh[0], kh[1], kh[2], bh[0], bh[1], bh[2]};
    int nv = HES ? 16 : (ENERGY ? 10 : 3);
    for (int q = 0; q < nv; q++) {
        double t = ttc_wave_sum(vals[q]);
        if (lane == 0) red[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        double t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        int q = threadIdx.x;
        int slot = q < 3 ? EV_FSX + q : (q < 10 ? EV_BOND + (q - 3) : (q < 13 ? EV_KHX + (q - 10) : EV_BHX + (q - 13)));
        partials[(size_t)slot * maxb + bid] = t;
    }
}

struct GatherPack {
    int n;
    const double4 *pos4;
    const int *orig, *inv;
    const float *qr;
    GatherConst gc;
    const int *b_off, *b_part;
    const double *b_r0, *b_k;
    double4 *g4;
    double *partials;
    int maxb, do_pme, do_bonds;
    const float *theta_near;
    const int *a_off, *a_ref;    // angle-type term forces per site (amask: kinds, bit k = kind k)
    const double4 *a_tf;
    int amask;
};

template <int P, bool ENERGY, bool HES = false>
__global__ void __launch_bounds__(256) k_gather(GatherPack a) {
    gather_body<P, ENERGY, HES>(blockIdx.x, a.n, a.pos4, a.orig, a.inv, a.qr, a.gc, a.b_off, a.b_part, a.b_r0, a.b_k,
                                a.g4, a.partials, a.maxb, a.do_pme, a.do_bonds, a.theta_near, a.a_off, a.a_ref,
                                a.a_tf, a.amask);
}
template <int P, bool ENERGY>
__global__ void __launch_bounds__(256) kb_gather(const GatherPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const GatherPack a = pk[m.rep];
    gather_body<P, ENERGY, false>(m.blk, a.n, a.pos4, a.orig, a.inv, a.qr, a.gc, a.b_off, a.b_part, a.b_r0, a.b_k,
                                  a.g4, a.partials, a.maxb, a.do_pme, a.do_bonds, a.theta_near, a.a_off, a.a_ref,
                                  a.a_tf, a.amask);
}

FftArgs fft_args(const ttc_ctx *c, int axis) {
    FftArgs f;
    f.K0 = c->K[0]; f.K1 = c->K[1]; f.K2 = c->K[2];
    const FftPlan &pl = c->plan[axis];
    f.n = pl.n;
    f.nrad = pl.nrad;
    int ns = 1;
    for (int i = 0; i < 16; i++) {
        f.rad[i] = i < pl.nrad ? pl.rad[i] : 1;
        f.ns[i] = ns;
        f.nr[i] = pl.n / f.rad[i];
        f.mag[i] = (unsigned)((0x100000000ull + ns - 1) / ns);   // 0 for ns = 1
        ns *= f.rad[i];
    }
    f.mag_n = (unsigned)((0x100000000ull + pl.n - 1) / pl.n);
    f.H = c->K[2] / 2 + 1;
    f.mag_h = (unsigned)((0x100000000ull + f.H - 1) / f.H);
    f.tw = pl.tw;
    return f;
}

}  // namespace

// ---------------------------------------------------------------- host side

// pme_setup.c:86-130 (bsp_mod) and the influence function of pme_calc.c:
// 52-72, evaluated once per box in FP64 and stored as FP32.
static void influence_host(const ttc_ctx *c, std::vector<float> &et) {
    int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2], P = c->order;
    // fillspline at w = 0 in FP64 (pme_spline.c:10-49)
    std::vector<double> a(P + 2, 0.0);
    {
        double w = 0.0;
        a[2] = w; a[1] = 1.0 - w; a[0] = 0.0;
        for (int i = 3; i <= P - 1; i++) {
            double div = 1.0 / (i - 1);
            a[i] = div * w * a[i - 1];
            for (int j = 1; j <= i - 2; j++) a[i - j] = div * ((w + j) * a[i - j - 1] + (i - j - w) * a[i - j]);
            a[1] *= div * (1.0 - w);
        }
        double div = 1.0 / (P - 1);
        a[P] = div * w * a[P - 1];
        for (int i = 1; i <= P - 2; i++) a[P - i] = div * ((w + i) * a[P - i - 1] + (P - i - w) * a[P - i]);
        a[1] *= div * (1.0 - w);
    }
    auto dftmod = [&](int K, std::vector<double> &bm) {
        std::vector<double> barr(K + P + 3, 0.0);
        for (int i = 2; i <= P + 1; i++) barr[i] = a[i - 1];
        bm.assign(K + 2, 0.0);
        for (int i = 1; i <= K; i++) {
            double s1 = 0.0, s2 = 0.0;
            for (int j = 1; j <= K; j++) {
                double arg = 2.0 * TTC_PI * (i - 1) * (j - 1) / K;
                s1 += barr[j] * cos(arg);
                s2 += barr[j] * sin(arg);
            }
            bm[i] = s1 * s1 + s2 * s2;
        }
    };
    std::vector<double> bx, by, bz;
    dftmod(K0, bx); dftmod(K1, by); dftmod(K2, bz);
    double lx = c->p.lx, ly = c->p.ly, lz = c->p.lz, vol = lx * ly * lz;
    double fac = TTC_PI * TTC_PI / c->p.kappa / c->p.kappa;
    // the half grid of the transforms: kz = 0 .. K2/2
    const int H = K2 / 2 + 1;
    et.assign((size_t)K0 * K1 * H, 0.0f);
    for (int k1 = 0; k1 < K0; k1++)
        for (int k2 = 0; k2 < K1; k2++)
            for (int k3 = 0; k3 < H; k3++) {
                size_t ind = ((size_t)k1 * K1 + k2) * H + k3;
                if (ind == 0) continue;   // m = 0 carries no energy; its value drops out of the force
                int mx = k1 - (k1 > K0 / 2 ? K0 : 0);
                int my = k2 - (k2 > K1 / 2 ? K1 : 0);
                int mz = k3 - (k3 > K2 / 2 ? K2 : 0);
                double hx = mx / lx, hy = my / ly, hz = mz / lz;
                double m2 = hx * hx + hy * hy + hz * hz;
                double bb = bx[k1 + 1] * by[k2 + 1] * bz[k3 + 1];
                et[ind] = (float)(exp(-fac * m2) / (TTC_PI * vol * m2 * bb));
            }
}

static int make_plan(FftPlan &pl, int n) {
    pl.n = n;
    pl.nrad = 0;
    int m = n;
    while (m % 4 == 0 && m > 4) { pl.rad[pl.nrad++] = 4; m /= 4; }
    while (m % 2 == 0) { pl.rad[pl.nrad++] = 2; m /= 2; }
    while (m % 3 == 0) { pl.rad[pl.nrad++] = 3; m /= 3; }
    for (int p = 5; m > 1; p++)
        while (m % p == 0) { pl.rad[pl.nrad++] = p; m /= p; }
    if (pl.nrad > 16) return TTC_ERR_ARG;
    std::vector<float2> tw(n);
    for (int i = 0; i < n; i++) tw[i] = make_float2((float)cos(2.0 * TTC_PI * i / n), (float)sin(2.0 * TTC_PI * i / n));
    if (hipMalloc(&pl.tw, sizeof(float2) * n) != hipSuccess) return TTC_ERR_ALLOC;
    if (hipMemcpy(pl.tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess) return TTC_ERR_HIP;
    return TTC_OK;
}

static int set_lds_attrs(ttc_ctx *c);

int ttc_pme_setup(ttc_ctx *c) {
    c->K[0] = c->p.gx; c->K[1] = c->p.gy; c->K[2] = c->p.gz;
    c->ngrid = (size_t)c->K[0] * c->K[1] * c->K[2];
    for (int a = 0; a < 3; a++) {
        // one FFT workgroup holds two buffers of FFT_LB lines of K + 1
        // complex values and K twiddles (8 (2 FFT_LB (K + 1) + K) bytes)
        // and its static 224-byte reduction array, far inside the default
        // 64 KiB of LDS; grids are limited to 246 points per axis (the
        // sizes tested reach 224)
        if (c->K[a] > 246) { ttc_set_error("PME grid dimension above 246 is not supported"); return TTC_ERR_ARG; }
        int r = make_plan(c->plan[a], c->K[a]);
        if (r) { ttc_set_error("PME grid size cannot be planned"); return r; }
    }
    TTC_HIP(hipMalloc(&c->qfix, sizeof(long long) * c->ngrid));
    TTC_HIP(hipMemset(c->qfix, 0, sizeof(long long) * c->ngrid));
    TTC_HIP(hipMalloc(&c->qc, sizeof(float2) * c->ngrid));
    TTC_HIP(hipMalloc(&c->qr, sizeof(float) * c->ngrid));
    TTC_HIP(hipMalloc(&c->eterm, sizeof(float) * c->ngrid));
    std::vector<float> et;
    influence_host(c, et);
    TTC_HIP(hipMemcpy(c->eterm, et.data(), sizeof(float) * et.size(), hipMemcpyHostToDevice));
    return TTC_OK;
}

int ttc_spread_attrs(ttc_ctx *c) { return set_lds_attrs(c); }
int ttc_pme_attrs(ttc_ctx *c) { return set_lds_attrs(c);
}

void ttc_pme_free(ttc_ctx *c) {
    hipFree(c->qfix); hipFree(c->qc); hipFree(c->qr); hipFree(c->eterm);
    if (c->theta_near) hipFree(c->theta_near);
    c->theta_near = nullptr;
    for (int a = 0; a < 3; a++) hipFree(c->plan[a].tw);
}

int ttc_pme_conv_blocks(ttc_ctx *c) { return (c->K[1] * (c->K[2] / 2 + 1) + FFT_LB - 1) / FFT_LB; }
int ttc_gather_blocks(ttc_ctx *c) { return (c->n + 255) / 256; }

static Spread spread_args(const ttc_ctx *c) {
    Spread sp;
    sp.lx = c->p.lx; sp.ly = c->p.ly; sp.lz = c->p.lz;
    sp.K0 = c->K[0]; sp.K1 = c->K[1]; sp.K2 = c->K[2];
    return sp;
}

static size_t brick_shm(const ttc_ctx *c, const Brick &br);

// workgroups of the spread that fit one CU's LDS (static part: mean force
// and a margin); at most 32 waves
static int spread_occupancy(size_t bytes) {
    return (int)std::min<size_t>(2048 / SPREAD_THREADS, (160 * 1024) / (bytes + 64));
}

// Bricks of 2x2x2 cells; on grids so fine that such a brick's tile (with
// the weights of 64 sites) does not fit one CU's LDS, the first of 2x2x1,
// 2x1x1 and 1x1x1 cells that fits.
static Brick brick_args(const ttc_ctx *c) {
    static const int shapes[4][3] = {{2, 2, 2}, {2, 2, 1}, {2, 1, 1}, {1, 1, 1}};
    Brick br;
    for (const auto &sh : shapes) {
        br.ncx = c->ncx; br.ncy = c->ncy; br.ncz = c->ncz;
        br.bx = c->ncx >= sh[0] ? sh[0] : 1; br.by = c->ncy >= sh[1] ? sh[1] : 1; br.bz = c->ncz >= sh[2] ? sh[2] : 1;
        br.nbx = (c->ncx + br.bx - 1) / br.bx; br.nby = (c->ncy + br.by - 1) / br.by;
        br.nbz = (c->ncz + br.bz - 1) / br.bz;
        br.cwx = c->p.lx / c->ncx; br.cwy = c->p.ly / c->ncy; br.cwz = c->p.lz / c->ncz;
        // the trigger fires at (rclist-rc)/4; one more step moves a site far less
        br.delta = 0.25 * (c->p.rclist - c->p.rc) + 0.25;
        br.chunk = 0;
        br.rec_off = 0;
        if (spread_occupancy(brick_shm(c, br) + 64 * SPREAD_REC_BYTES(c->order)) >= 1) break;
    }
    return br;
}

// the brick with its LDS layout: tile, then the weights of `chunk` sites,
// chunk the largest of 256/128/64 that keeps the occupancy of chunk 64
// (ttc_tune may set the chunk instead: c->spread_chunk, kept only while it
// fits the LDS)
static Brick spread_brick(const ttc_ctx *c) {
    Brick br = brick_args(c);
    size_t tile = brick_shm(c, br);
    br.rec_off = (int)(tile / 8);
    br.chunk = 64;
    if (c->spread_chunk > 0 && spread_occupancy(tile + (size_t)c->spread_chunk * SPREAD_REC_BYTES(c->order)) >= 1) {
        br.chunk = c->spread_chunk;
        return br;
    }
    const int occ64 = spread_occupancy(tile + 64 * SPREAD_REC_BYTES(c->order));
    for (int ch : {256, 128})
        if (spread_occupancy(tile + (size_t)ch * SPREAD_REC_BYTES(c->order)) >= occ64) { br.chunk = ch; break; }
    return br;
}

// the chunks the tuner may try (64, 128, 256, those that fit the LDS)
int ttc_spread_chunk_ok(const ttc_ctx *c, int chunk) {
    Brick br = brick_args(c);
    size_t tile = brick_shm(c, br);
    return spread_occupancy(tile + (size_t)chunk * SPREAD_REC_BYTES(c->order)) >= 1;
}

// dynamic LDS of the spread: tile + weights
static size_t spread_shm(const ttc_ctx *c, const Brick &br) {
    return (size_t)br.rec_off * 8 + (size_t)br.chunk * SPREAD_REC_BYTES(c->order);
}

// exact largest tile over all bricks, by the kernel's own formula
static size_t brick_shm(const ttc_ctx *c, const Brick &br) {
    auto ext = [&](int nc, int bc, double cw, double l, int k) {
        int mx = 0;
        for (int c0 = 0; c0 < nc; c0 += bc) {
            int c1 = std::min(c0 + bc, nc);
            double xlo = -0.5 * l + c0 * cw - br.delta, xhi = -0.5 * l + c1 * cw + br.delta;
            int glo = (int)floor((double)k * (xlo / l + 0.5));
            int ghi = (int)floor((double)k * (xhi / l + 0.5));
            mx = std::max(mx, ghi - glo + c->order);
        }
        return mx;
    };
    size_t t = (size_t)ext(br.ncx, br.bx, br.cwx, c->p.lx, c->K[0]) * ext(br.ncy, br.by, br.cwy, c->p.ly, c->K[1]) *
               ext(br.ncz, br.bz, br.cwz, c->p.lz, c->K[2]);
    return t * sizeof(unsigned long long);
}

// the spread of the current positions (sorted order, the cell ranges of
// the last rebuild)
static SpreadPack spread_pack(ttc_ctx *c) {
    return SpreadPack{c->pos4, spread_args(c), spread_brick(c), c->cell_start, c->cell_cnt, c->qfix};
}

static int spread_blocks(const ttc_ctx *c) {
    Brick br = brick_args(c);
    return br.nbx * br.nby * br.nbz;
}

static void launch_spread_pack(ttc_ctx *c, const SpreadPack &a) {
    const size_t shm = spread_shm(c, a.br);
    if (c->order == 4) k_spread<4><<<spread_blocks(c), SPREAD_THREADS, shm, c->stream>>>(a);
    else k_spread<6><<<spread_blocks(c), SPREAD_THREADS, shm, c->stream>>>(a);
}

// spread only (ttc_pme_calc / ttc_cnonbondnl, the redo of an overflowed step)
void ttc_launch_spread(ttc_ctx *c) { launch_spread_pack(c, spread_pack(c)); }

// the drift alone with the spread's copies and the drifted positions'
// fixed-point coordinates (which the half-list pair kernel reads); the
// velocity-Verlet loop's (vv) also writes the original-order snapshot and
// takes the gather's output in the original numbering
static MdDriftPack md_drift_pack(ttc_ctx *c, double vscal, bool vv) {
    MdDriftPack a{c->pos4, c->vel4, c->f4, c->p.dt, vscal, spread_args(c), c->n, c->ncell, c->cell_start,
                  c->cell_cnt, c->scs, c->scc, c->spos, c->status, c->anchor, c->ev + EV_N,
                  KickIn{c->fpair, c->g4, c->partials, c->nblk_max, ttc_gather_blocks(c), c->n, vscal, nullptr,
                         nullptr, c->fixj}};
    a.ki.posq = c->posq;
    a.ki.qkx = 4294967296.0 / c->p.lx;
    a.ki.qky = 4294967296.0 / c->p.ly;
    a.ki.qkz = 4294967296.0 / c->p.lz;
    if (vv) {
        a.ki.g4 = c->g4o;
        a.ki.orig = c->orig;
        a.ki.snap = c->snap;
    }
    return a;
}

static int md_drift_blocks(const ttc_ctx *c) { return (std::max(c->n, c->ncell) + 255) / 256; }

void ttc_launch_md_drift(ttc_ctx *c, double vscal, bool kick) {
    const MdDriftPack a = md_drift_pack(c, vscal, true);
    if (kick) k_md_drift<true><<<md_drift_blocks(c), 256, 0, c->stream>>>(a);
    else k_md_drift<false><<<md_drift_blocks(c), 256, 0, c->stream>>>(a);
}

void ttc_launch_nhc_drift(ttc_ctx *c) {
    k_md_drift<false><<<md_drift_blocks(c), 256, 0, c->stream>>>(md_drift_pack(c, 1.0, false));
}

// the spread from the drift's copies
static SpreadPack md_spread_pack(ttc_ctx *c) {
    SpreadPack a = spread_pack(c);
    a.pos4 = c->spos;
    a.cell_start = c->scs;
    a.cell_cnt = c->scc;
    return a;
}

void ttc_launch_md_spread(ttc_ctx *c) { launch_spread_pack(c, md_spread_pack(c)); }

int ttc_spread_shm(const ttc_ctx *c) { return (int)spread_shm(c, spread_brick(c)); }

// allow the spread tile and the FFT tiles above 64 KiB of LDS (gfx950 has
// 160 KiB per CU).  The attribute belongs to the kernel, not the context:
// it only grows, so a context with a smaller tile (or a tuned chunk) never
// lowers it below what another context launches with.
static int set_lds_attrs(ttc_ctx *c) {
    static int attr_max[64] = {0};   // per device
    int shm = (int)spread_shm(c, spread_brick(c));
    if (spread_occupancy((size_t)shm) < 1) {
        ttc_set_error("PME spread tile exceeds the LDS of one CU");
        return TTC_ERR_ARG;
    }
    int dev = 0;
    TTC_HIP(hipGetDevice(&dev));
    int &amax = attr_max[dev & 63];
    if (shm <= amax) return TTC_OK;
    amax = shm;
    const void *fns[] = {(const void *)k_spread<4>, (const void *)k_spread<6>};
    for (const void *f : fns) TTC_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    return TTC_OK;
}

static size_t fft_shm(const ttc_ctx *c, int axis) {
    // two buffers of FFT_LB lines (padded to n + 1) and the twiddles
    return sizeof(float2) * (2 * FFT_LB * (c->plan[axis].n + 1) + c->plan[axis].n);
}

// the five line passes of one step: z forward, y forward, x forward +
// influence function + x inverse, y inverse, z inverse
enum { FP_Z_FWD = 0, FP_Y_FWD, FP_X_CONV, FP_Y_INV, FP_Z_INV, FP_N };
static const int kPassAxis[FP_N] = {2, 1, 0, 1, 2};
static const int kPassSign[FP_N] = {+1, +1, +1, -1, -1};

static FftPack fft_pack(const ttc_ctx *c, int pass) {
    return FftPack{fft_args(c, kPassAxis[pass]), kPassSign[pass], c->qfix, c->qc, c->qr, c->eterm,
                   TTC_PI * TTC_PI / c->p.kappa / c->p.kappa, c->p.lx, c->p.ly, c->p.lz, c->partials, c->nblk_max};
}

static int fft_blocks(const ttc_ctx *c, int pass) {
    const int H = c->K[2] / 2 + 1;
    if (pass == FP_Z_FWD || pass == FP_Z_INV) return (c->K[0] * c->K[1] + 2 * FFT_LB - 1) / (2 * FFT_LB);
    if (pass == FP_X_CONV) return (c->K[1] * H + FFT_LB - 1) / FFT_LB;
    return (c->K[0] * H + FFT_LB - 1) / FFT_LB;
}

// FFT passes and influence function.  Expects the spread grid in qfix;
// leaves qfix zeroed.
void ttc_launch_fft(ttc_ctx *c, bool energy) {
#define FFTL(MODE, E, PASS) \
    k_fft<MODE, E><<<fft_blocks(c, PASS), FFT_THREADS, fft_shm(c, kPassAxis[PASS]), c->stream>>>(fft_pack(c, PASS))
    FFTL(0, false, FP_Z_FWD);
    FFTL(1, false, FP_Y_FWD);
    if (energy) FFTL(2, true, FP_X_CONV);
    else FFTL(2, false, FP_X_CONV);
    FFTL(1, false, FP_Y_INV);
    FFTL(3, false, FP_Z_INV);
#undef FFTL
}

// the gather's angle-kind mask (bit k = kind k) from a TTC_BT_* selection
static int angle_mask(const ttc_ctx *c, int bterms) { return c->ang_start[4] ? (bterms >> 1) & 15 : 0; }

void ttc_launch_gather(ttc_ctx *c, bool energy, bool pme, int bterms, bool hes) {
    int am = angle_mask(c, bterms);
    if (am) ttc_launch_angles(c, energy, hes, bterms, false);
    double lx = c->p.lx, ly = c->p.ly, lz = c->p.lz;
    int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2];
    GatherConst gc;
    gc.lx = lx; gc.ly = ly; gc.lz = lz;
    gc.hx = 0.5 * lx; gc.hy = 0.5 * ly; gc.hz = 0.5 * lz;
    gc.K0 = K0; gc.K1 = K1; gc.K2 = K2;
    gc.qscale = TTC_QFACTOR / c->p.eps_rf;
    int blocks = ttc_gather_blocks(c);
    GatherPack a{c->n, c->pos4, c->orig, c->inv, c->qr, gc, c->b_off, c->b_part, c->b_r0, c->b_k, c->g4,
                 c->partials, c->nblk_max, pme ? 1 : 0, (bterms & TTC_BT_BOND) ? 1 : 0, c->theta_near,
                 c->ang_off, c->ang_ref, c->ang_tf, am};
#define GATHER(P, E, H) k_gather<P, E, H><<<blocks, 256, 0, c->stream>>>(a)
    if (c->order == 4) {
        if (hes) GATHER(4, true, true); else if (energy) GATHER(4, true, false); else GATHER(4, false, false);
    } else {
        if (hes) GATHER(6, true, true); else if (energy) GATHER(6, true, false); else GATHER(6, false, false);
    }
#undef GATHER
}

namespace {
// snap[orig[s]] = pos4[s]: the snapshot an MD step's drift writes, from the
// current positions
__global__ void k_snap(int n, const double4 *__restrict__ pos4, const int *__restrict__ orig, double4 *snap) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) snap[orig[s]] = pos4[s];
}
}  // namespace

void ttc_launch_snap(ttc_ctx *c) { k_snap<<<(c->n + 255) / 256, 256, 0, c->stream>>>(c->n, c->pos4, c->orig, c->snap); }

void ttc_launch_gather_orig(ttc_ctx *c, bool energy) {
    int am = angle_mask(c, TTC_BT_ALL);
    if (am) ttc_launch_angles(c, energy, false, TTC_BT_ALL, true);
    double lx = c->p.lx, ly = c->p.ly, lz = c->p.lz;
    GatherConst gc;
    gc.lx = lx; gc.ly = ly; gc.lz = lz;
    gc.hx = 0.5 * lx; gc.hy = 0.5 * ly; gc.hz = 0.5 * lz;
    gc.K0 = c->K[0]; gc.K1 = c->K[1]; gc.K2 = c->K[2];
    gc.qscale = TTC_QFACTOR / c->p.eps_rf;
    int blocks = ttc_gather_blocks(c);
    GatherPack a{c->n, c->snap, nullptr, nullptr, c->qr, gc, c->b_off, c->b_part, c->b_r0, c->b_k, c->g4o,
                 c->partials, c->nblk_max, 1, 1, c->theta_near, c->ang_off, c->ang_ref, c->ang_tf, am};
#define GATHER(P, E) k_gather<P, E, false><<<blocks, 256, 0, c->stream>>>(a)
    if (c->order == 4) { if (energy) GATHER(4, true); else GATHER(4, false); }
    else { if (energy) GATHER(6, true); else GATHER(6, false); }
#undef GATHER
}

static void launch_pme_impl(ttc_ctx *c, bool energy, bool pme, int bterms, bool hes) {
    if (pme) ttc_launch_fft(c, energy);
    ttc_launch_gather(c, energy, pme, bterms, hes);
}

void ttc_launch_pme(ttc_ctx *c, bool energy, bool pme, int bterms) { launch_pme_impl(c, energy, pme, bterms, false); }

void ttc_launch_pme_hes(ttc_ctx *c, bool pme, int bterms) { launch_pme_impl(c, true, pme, bterms, true); }

// theta near the origin: the FFT/convolution pipeline run on a unit charge
// at grid point 0 leaves theta(m) in qr (the kernel the gather convolves
// with); offsets |d| < order per axis are kept.  Runs once per context.
int ttc_pme_theta(ttc_ctx *c) {
    if (c->theta_near) return TTC_OK;
    const int P = c->order, ND = 2 * P - 1;
    int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2];
    TTC_HIP(hipStreamSynchronize(c->stream));
    long long one = (long long)TTC_QFIX_SCALE;
    TTC_HIP(hipMemcpyAsync(c->qfix, &one, sizeof one, hipMemcpyHostToDevice, c->stream));
    launch_pme_impl(c, false, true, false, false);   // FFTs; the gather output is not used
    TTC_HIP(hipGetLastError());
    std::vector<float> grid((size_t)K0 * K1 * K2);
    TTC_HIP(hipMemcpyAsync(grid.data(), c->qr, sizeof(float) * grid.size(), hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> near((size_t)ND * ND * ND);
    for (int ox = 0; ox < ND; ox++)
        for (int oy = 0; oy < ND; oy++)
            for (int oz = 0; oz < ND; oz++) {
                int x = ((ox - (P - 1)) % K0 + K0) % K0, y = ((oy - (P - 1)) % K1 + K1) % K1,
                    z = ((oz - (P - 1)) % K2 + K2) % K2;
                near[(ox * ND + oy) * ND + oz] = grid[((size_t)x * K1 + y) * K2 + z];
            }
    TTC_HIP(hipMalloc(&c->theta_near, sizeof(float) * near.size()));
    TTC_HIP(hipMemcpy(c->theta_near, near.data(), sizeof(float) * near.size(), hipMemcpyHostToDevice));
    return TTC_OK;
}

// ---- batched stages of the NHC step (ttc_batch.hip) ------------------------

int ttc_bstage_drift(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<MdDriftPack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) {
        pk.push_back(md_drift_pack(c, 1.0, false));
        nb.push_back(md_drift_blocks(c));
    }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_drift(const TtcStage &st, hipStream_t s) {
    kb_md_drift<<<st.nblk, 256, 0, s>>>((const MdDriftPack *)st.pk, st.map);
}

int ttc_bstage_spread(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<SpreadPack> pk;
    std::vector<int> nb;
    size_t shm = 0;
    for (ttc_ctx *c : r) {
        if (c->order != 4) {
            ttc_set_error("ttc_batch: batched launches need spline order 4 in every replica");
            return TTC_ERR_ARG;
        }
        pk.push_back(md_spread_pack(c));
        nb.push_back(spread_blocks(c));
        shm = std::max(shm, spread_shm(c, pk.back().br));
    }
    if (spread_occupancy(shm) < 1) {
        ttc_set_error("ttc_batch: PME spread tile exceeds the LDS of one CU");
        return TTC_ERR_ARG;
    }
    TTC_HIP(hipFuncSetAttribute((const void *)kb_spread<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
    return ttc_stage_make(st, pk, nb, shm);
}
void ttc_blaunch_spread(const TtcStage &st, hipStream_t s) {
    kb_spread<4><<<st.nblk, SPREAD_THREADS, st.shm, s>>>((const SpreadPack *)st.pk, st.map);
}

// st[FP_N]: the five FFT passes
int ttc_bstage_fft(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    for (int pass = 0; pass < FP_N; pass++) {
        std::vector<FftPack> pk;
        std::vector<int> nb;
        size_t shm = 0;
        for (ttc_ctx *c : r) {
            pk.push_back(fft_pack(c, pass));
            nb.push_back(fft_blocks(c, pass));
            shm = std::max(shm, fft_shm(c, kPassAxis[pass]));
        }
        int e = ttc_stage_make(&st[pass], pk, nb, shm);
        if (e) return e;
    }
    return TTC_OK;
}
void ttc_blaunch_fft(const TtcStage *st, bool energy, hipStream_t s) {
#define FFTB(MODE, E, PASS) \
    kb_fft<MODE, E><<<st[PASS].nblk, FFT_THREADS, st[PASS].shm, s>>>((const FftPack *)st[PASS].pk, st[PASS].map)
    FFTB(0, false, FP_Z_FWD);
    FFTB(1, false, FP_Y_FWD);
    if (energy) FFTB(2, true, FP_X_CONV);
    else FFTB(2, false, FP_X_CONV);
    FFTB(1, false, FP_Y_INV);
    FFTB(3, false, FP_Z_INV);
#undef FFTB
}

int ttc_bstage_gather(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<GatherPack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) {
        GatherConst gc;
        gc.lx = c->p.lx; gc.ly = c->p.ly; gc.lz = c->p.lz;
        gc.hx = 0.5 * c->p.lx; gc.hy = 0.5 * c->p.ly; gc.hz = 0.5 * c->p.lz;
        gc.K0 = c->K[0]; gc.K1 = c->K[1]; gc.K2 = c->K[2];
        gc.qscale = TTC_QFACTOR / c->p.eps_rf;
        pk.push_back(GatherPack{c->n, c->pos4, c->orig, c->inv, c->qr, gc, c->b_off, c->b_part, c->b_r0, c->b_k, c->g4,
                                c->partials, c->nblk_max, 1, 1, c->theta_near, c->ang_off, c->ang_ref, c->ang_tf,
                                angle_mask(c, TTC_BT_ALL)});
        nb.push_back(ttc_gather_blocks(c));
    }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_gather(const TtcStage &st, bool energy, hipStream_t s) {
    if (energy) kb_gather<4, true><<<st.nblk, 256, 0, s>>>((const GatherPack *)st.pk, st.map);
    else kb_gather<4, false><<<st.nblk, 256, 0, s>>>((const GatherPack *)st.pk, st.map);
}
