// ttc_nblist.hip -- neighbour list on the device (replaces nblist.c).
//
// Rebuild = three kernels, each a no-op unless the displacement trigger is
// set (so the whole sequence can sit in a captured HIP graph every step):
//   R1 cell_assign  : cell of every site, counted into fixed-capacity cells;
//                     the workgroup that finishes last scans the counts into
//                     cell starts
//   R2 reorder      : rank inside the cell by original index (deterministic
//                     order without a sort), scatter the site state into the
//                     B buffers at its new sorted position
//   R3 build        : copy B -> A, anchors, and the full Verlet list: the
//                     reference's FP64 test for explicit builds (k_build),
//                     the group-cooperative FP32 build of the pruned list
//                     inside MD loops (k_build_grp)
// Cells are at least rclist/2 wide, so a 5x5x5 stencil covers rclist.
#include "ttc_common.h"

#include <algorithm>
#include <type_traits>

namespace {

struct CellGeom {
    int ncx, ncy, ncz;
    double lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ int cell_coord(double x, double h, double l, int nc) {
    int c = (int)((x + h) * (double)nc / l);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// a conditional (MD-loop) rebuild runs only when the trigger fired, and not
// after an earlier rebuild of the same loop overflowed (the host grows the
// list and redoes that step)
__device__ __forceinline__ bool rebuild_off(const int *status, int conditional) {
    return conditional && (status[ST_NLFLAG] == 0 || (status[ST_OVERFLOW] & TTC_OVF_REDO) != 0);
}

// Counts to cell starts in one workgroup of 256 threads: chunks of
// 256 x 16 cells are read coalesced at device scope (the other workgroups'
// atomics) into LDS, each thread scans a run of 16, and the starts are
// written back coalesced; cell_acc is reset to zero on the way.
#define SCAN_PER 16
__device__ void cell_scan_block(int ncell, int *cell_acc, int *cell_cnt, int *cell_start, int *status) {
    __shared__ int sc[256 * SCAN_PER];
    __shared__ int wsum[4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int carry = 0;
    for (int base = 0; base < ncell; base += 256 * SCAN_PER) {
        int vv[SCAN_PER];
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++) {   // all loads in flight at once
            const int c = base + k * 256 + t;
            vv[k] = c < ncell ? __hip_atomic_load(&cell_acc[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        }
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++) {
            const int c = base + k * 256 + t;
            if (c < ncell) {
                cell_cnt[c] = vv[k];
                cell_acc[c] = 0;
            }
            sc[k * 256 + t] = vv[k];
        }
        __syncthreads();
        int sum = 0;
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++) sum += sc[t * SCAN_PER + k];
        int incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int run = carry + incl - sum;
        for (int w = 0; w < wave; w++) run += wsum[w];
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++) {
            const int v = sc[t * SCAN_PER + k];
            sc[t * SCAN_PER + k] = run;
            run += v;
        }
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SCAN_PER; k++) {
            const int c = base + k * 256 + t;
            if (c < ncell) cell_start[c] = sc[k * 256 + t];
        }
        __syncthreads();
    }
    if (t == 0) {
        status[ST_NREBUILD] += 1;
        status[ST_TICKET] = 0;   // ready for the next rebuild
    }
}

// Cell binning of every site; the workgroup that finishes last (atomic
// ticket, no waiting) turns the counts into cell starts.
__device__ __forceinline__ void cell_assign_body(int bid, int nblk, int n, const double4 *__restrict__ pos4,
                                                 const CellGeom &g, int cellcap, int ncell, int *cell_acc,
                                                 int *cell_cnt, int *cell_start, int *cell_of, int *members,
                                                 int *status, int conditional) {
    if (rebuild_off(status, conditional)) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i < n) {
        double4 p = pos4[i];
        int cx = cell_coord(p.x, g.hx, g.lx, g.ncx);
        int cy = cell_coord(p.y, g.hy, g.ly, g.ncy);
        int cz = cell_coord(p.z, g.hz, g.lz, g.ncz);
        int c = cx + g.ncx * (cy + g.ncy * cz);
        int slot = atomicAdd(&cell_acc[c], 1);
        cell_of[i] = c;
        if (slot < cellcap) members[(size_t)c * cellcap + slot] = i;
        else atomicOr(&status[ST_OVERFLOW], 2);
    }
    // Every count is a device-scope atomic whose old value this workgroup
    // has already consumed (the slot), so the counts are performed before
    // the ticket; the last workgroup reads them with device-scope loads.  No
    // fence is needed (an agent-scope release would write back the L2).
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&status[ST_TICKET], 1) == nblk - 1;
    __syncthreads();
    if (!last) return;
    cell_scan_block(ncell, cell_acc, cell_cnt, cell_start, status);
}

struct CellAssignPack {
    int n, nblk;
    const double4 *pos4;
    CellGeom g;
    int cellcap, ncell;
    int *cell_acc, *cell_cnt, *cell_start, *cell_of, *members, *status;
    int conditional;
};

__device__ __forceinline__ void cell_assign_pack_body(int bid, const CellAssignPack &a) {
    cell_assign_body(bid, a.nblk, a.n, a.pos4, a.g, a.cellcap, a.ncell, a.cell_acc, a.cell_cnt, a.cell_start,
                     a.cell_of, a.members, a.status, a.conditional);
}
__global__ void __launch_bounds__(256) k_cell_assign(CellAssignPack a) { cell_assign_pack_body(blockIdx.x, a); }
__global__ void __launch_bounds__(256) kb_cell_assign(const CellAssignPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    cell_assign_pack_body(m.blk, pk[m.rep]);
}

struct ReorderPack {
    int n, cellcap;
    const int *cell_of, *cell_cnt, *cell_start, *members, *orig;
    const double4 *pos4, *vel4, *f4;
    const int *type;
    double4 *pos4b, *vel4b, *f4b;
    int *typeb, *origb;
    const double4 *fs, *fl;
    double4 *fsb, *flb;
    float4 *posfb;
    int *inv, *status;
    int conditional;
};

__device__ __forceinline__ void reorder_body(int bid, const ReorderPack &a) {
    const int n = a.n, cellcap = a.cellcap;
    const int *cell_of = a.cell_of, *cell_cnt = a.cell_cnt, *cell_start = a.cell_start, *members = a.members,
              *orig = a.orig, *type = a.type;
    const double4 *pos4 = a.pos4, *vel4 = a.vel4, *f4 = a.f4, *fs = a.fs, *fl = a.fl;
    double4 *pos4b = a.pos4b, *vel4b = a.vel4b, *f4b = a.f4b, *fsb = a.fsb, *flb = a.flb;
    int *typeb = a.typeb, *origb = a.origb, *inv = a.inv;
    float4 *posfb = a.posfb;
    if (rebuild_off(a.status, a.conditional)) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = cell_of[i];
    int m = min(cell_cnt[c], cellcap);
    int o = orig[i];
    int r = 0;
    const int *mem = members + (size_t)c * cellcap;
    for (int k = 0; k < m; k++) r += (orig[mem[k]] < o);
    int s = cell_start[c] + r;
    double4 p = pos4[i];
    pos4b[s] = p;
    // candidates of the fast build: FP32 position, original index | LJ type << TTC_TSHIFT
    posfb[s] = make_float4((float)p.x, (float)p.y, (float)p.z, __int_as_float(o | (type[i] << TTC_TSHIFT)));
    vel4b[s] = vel4[i];
    f4b[s] = f4[i];   // the last force feeds the next half kick
    fsb[s] = fs[i];   // RESPA short / long forces likewise
    flb[s] = fl[i];
    typeb[s] = type[i];
    origb[s] = o;
    inv[o] = s;
}

__global__ void k_reorder(ReorderPack a) { reorder_body(blockIdx.x, a); }
__global__ void kb_reorder(const ReorderPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const ReorderPack a = pk[m.rep];
    reorder_body(m.blk, a);
}

#define BUILD_WAVES 8

// The reference's pair test (nblist.c:116-145): |d| wrapped once, every
// component <= rcut, dr2 < rcut2 in IEEE double rounded op by op.  It is
// symmetric in the two sites, so the full list holds every reference pair in
// both rows.  Evaluated without branches.
__device__ __forceinline__ bool ref_pair_test(const double4 &pi, const double4 &pj, const CellGeom &g,
                                              double rcut, double rcut2) {
    double dx = fabs(pi.x - pj.x);
    dx = dx > g.hx ? dx - g.lx : dx;
    double dy = fabs(pi.y - pj.y);
    dy = dy > g.hy ? dy - g.ly : dy;
    double dz = fabs(pi.z - pj.z);
    dz = dz > g.hz ? dz - g.lz : dz;
    double dr2 = ttc_r2_exact(dx, dy, dz);
    return (dx <= rcut) & (dy <= rcut) & (dz <= rcut) & (dr2 < rcut2);
}

// smallest periodic distance from coordinate x to the slab [lo, lo + w)
__device__ __forceinline__ double slab_gap(double x, double lo, double w, double l) {
    double d = x - lo;
    d -= l * floor(d / l);           // 0 <= d < l
    if (d < w) return 0.0;
    return fmin(d - w, l - d);
}

// Full Verlet list.  One workgroup per ELL group of 64 sorted sites; its
// BUILD_WAVES waves split each site's 5x5x5 cell stencil (stencil cell t
// goes to wave t % BUILD_WAVES) and skip cells farther than rcut.  Pass 1
// counts each site's hits per wave, pass 2 writes them at prefix offsets,
// so the row order is fixed (stencil order within a wave, waves in order).
// This is the reference's list (nblist() of the C ABI, exports, lists at
// the anchors); lists of MD loops come from k_build_grp.
__global__ void __launch_bounds__(64 * BUILD_WAVES)
k_build(int n, int cap, CellGeom g, double rcut, double rcut2, const double4 *__restrict__ pos4b,
        const double4 *__restrict__ vel4b, const double4 *__restrict__ f4b, const int *__restrict__ typeb,
        const int *__restrict__ origb, double4 *pos4, double4 *vel4, double4 *f4, int *type, int *orig,
        double4 *anchor, int *inv, const double4 *__restrict__ fsb, const double4 *__restrict__ flb,
        double4 *fs, double4 *fl,
        const int *__restrict__ cell_cnt, const int *__restrict__ cell_start,
        const int *__restrict__ ex_off, const int *__restrict__ ex_part, const int *__restrict__ ex_flag,
        int ex_span, uint32_t *nbl, int *glen, int *gcnt, int *status, int conditional, int copy_state) {
    if (rebuild_off(status, conditional)) return;
    __shared__ int scnt[BUILD_WAVES][64];
    __shared__ int smax;
    // hit masks of the (at most 125) stencil cells: cell t of a site goes to
    // wave t % BUILD_WAVES, slot t / BUILD_WAVES
    __shared__ unsigned long long hm[BUILD_WAVES][(125 + BUILD_WAVES - 1) / BUILD_WAVES][64];
    int grp = blockIdx.x;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int s = grp * 64 + lane;
    bool valid = s < n;
    double4 pi = valid ? pos4b[s] : make_double4(0, 0, 0, 0);
    int oi = valid ? origb[s] : 0;
    if (valid && copy_state && wave == 0) {
        pos4[s] = pi;
        vel4[s] = vel4b[s];
        f4[s] = f4b[s];
        fs[s] = fsb[s];
        fl[s] = flb[s];
        type[s] = typeb[s];
        orig[s] = oi;
        anchor[s] = pi;
        inv[oi] = s;
    }
    int cx = cell_coord(pi.x, g.hx, g.lx, g.ncx);
    int cy = cell_coord(pi.y, g.hy, g.ly, g.ncy);
    int cz = cell_coord(pi.z, g.hz, g.lz, g.ncz);
    int nsx = g.ncx >= 5 ? 5 : g.ncx, nsy = g.ncy >= 5 ? 5 : g.ncy, nsz = g.ncz >= 5 ? 5 : g.ncz;
    int nst = nsx * nsy * nsz;
    double cwx = g.lx / g.ncx, cwy = g.ly / g.ncy, cwz = g.lz / g.ncz;
    int xb = valid ? ex_off[oi] : 0, xe = valid ? ex_off[oi + 1] : 0;
    uint32_t *row = nbl + (size_t)grp * cap * 64;

    // pass 0 tests every candidate once and keeps, per stencil cell, the hits
    // among its first 64 sites as a bit mask in LDS; pass 1 writes the
    // entries from the masks (re-testing only sites beyond the 64th of a
    // cell), at the per-wave prefix offsets of the pass-0 counts
    int mine = 0, off = 0;
    for (int pass = 0; pass < 2; pass++) {
        int k = off;
        for (int t = wave, slot = 0; t < nst && valid; t += BUILD_WAVES, slot++) {
            int ox = t % nsx, oy = (t / nsx) % nsy, oz = t / (nsx * nsy);
            int x = g.ncx >= 5 ? (cx + ox - 2 + g.ncx) % g.ncx : ox;
            int y = g.ncy >= 5 ? (cy + oy - 2 + g.ncy) % g.ncy : oy;
            int z = g.ncz >= 5 ? (cz + oz - 2 + g.ncz) % g.ncz : oz;
            double gx = slab_gap(pi.x, -g.hx + x * cwx, cwx, g.lx);
            double gy = slab_gap(pi.y, -g.hy + y * cwy, cwy, g.ly);
            double gz = slab_gap(pi.z, -g.hz + z * cwz, cwz, g.lz);
            // margin: the test is exact, the pruning only needs to be safe
            if (gx * gx + gy * gy + gz * gz > rcut2 * 1.0001 + 1e-9) continue;
            int c = x + g.ncx * (y + g.ncy * z);
            int j0 = cell_start[c], j1 = j0 + cell_cnt[c];
            unsigned long long mask = 0ull;
            int jt = j0;   // first site tested again in pass 1
            if (pass == 0) {
                for (int j = j0; j < j1; j++) {
                    if (j == s || !ref_pair_test(pi, pos4b[j], g, rcut, rcut2)) continue;
                    mine++;
                    if (j - j0 < 64) mask |= 1ull << (j - j0);
                }
                hm[wave][slot][lane] = mask;
                continue;
            }
            mask = hm[wave][slot][lane];
            jt = j0 + 64;
            while (true) {
                int j;
                if (mask) {
                    j = j0 + __builtin_ctzll(mask);
                    mask &= mask - 1;
                } else {
                    // sites past the 64th of the cell (dense cells only)
                    while (jt < j1 && (jt == s || !ref_pair_test(pi, pos4b[jt], g, rcut, rcut2)))
                        jt++;
                    if (jt >= j1) break;
                    j = jt++;
                }
                uint32_t cls = 0;
                int oj = origb[j];
                int d = oj - oi;
                if (d <= ex_span && -d <= ex_span) {
                    for (int q = xb; q < xe; q++)
                        if (ex_part[q] == oj) { int f = ex_flag[q]; cls = f == 4 ? 2u : (f ? 1u : 0u); }
                }
                if (k < cap) row[ttc_nbl_idx(k, lane)] = (uint32_t)j | ((uint32_t)typeb[j] << TTC_TSHIFT) | (cls << 30);
                k++;
            }
        }
        if (pass == 0) {
            scnt[wave][lane] = mine;
            __syncthreads();
            for (int w = 0; w < wave; w++) off += scnt[w][lane];
        }
    }
    // row length of the group = longest row; pad the shorter ones
    int total = 0;
    for (int w = 0; w < BUILD_WAVES; w++) total += scnt[w][lane];
    if (wave == 0) {
        int full = total, sum = total;
        for (int o = 32; o > 0; o >>= 1) {
            full = max(full, __shfl_xor(full, o, 64));
            sum += __shfl_xor(sum, o, 64);
        }
        if (lane == 0) {
            smax = full;
            gcnt[grp] = sum;
            glen[grp] = min(full, cap);
            if (full > cap) atomicOr(&status[ST_OVERFLOW], 1);
            atomicMax(&status[ST_MAXROW], full);
        }
    }
    __syncthreads();
    int m = min((smax + 15) & ~15, cap);   // whole 16-entry tiles
    for (int k = min(total, cap) + wave; k < m; k += BUILD_WAVES) row[ttc_nbl_idx(k, lane)] = TTC_PAD;
}

// ---- list of MD loops: group-cooperative build ------------------------------
//
// The list rebuilt inside MD loops only has to be a superset of the pairs
// within the exact cutoff (the pair kernel applies the FP64 test to every
// entry), so it is built from the FP32 positions at a cutoff widened by a
// margin.  One workgroup per ELL group of 64 sorted sites, GB_WAVES waves of
// GB_ROWS rows (sites) each.  Every wave works on its own rows alone:
//   1. it takes its rows' bounding box in the minimum-image frame of its
//      first row (compact even where the rows straddle the periodic
//      boundary) and enumerates the rows of cells (fixed y, z cell) that come
//      within the cutoff of it, one cell row per lane.  Sites are sorted
//      x-fastest by cell, so each cell row contributes one or two contiguous
//      ranges of sorted sites, each with one periodic shift into the frame
//      (or, in boxes where the range wraps onto itself, the minimum-image
//      flag).
//   2. it walks the ranges in lane order, 64 consecutive candidate sites per
//      load (lane = candidate, one coalesced 16-byte load each), and tests
//      each of its rows against the 64 candidates with the row's coordinates
//      wave-uniform; one ballot gives the row's hits, and each hit lane
//      writes its entry at the row's running count plus the number of hits
//      below it.
// The row order is therefore fixed (range order, then sorted order), no
// cross-wave exchange is needed and nothing limits the number of ranges.
// Small per-wave boxes keep the candidate count near twice the sites inside
// the cutoff sphere's bounding cells rather than the group's.
#define GB_WAVES 4
#define GB_ROWS (64 / GB_WAVES)
#define GB_MI (1 << 13)
#define GB_QCAP 128   // candidate ring per wave (two batches of 64)
#define GB_EXMAX 8    // exclusion partners per site held in LDS (more: read from the CSR)
#define GB_XM_SPAN 15 // original-index offsets +-15 fit a 2-bit class mask of 62 bits
#define GB_ZMARGIN 1e-3f   // A: far above FP32 rounding of frame coordinates

struct GbGeom {
    int ncx, ncy, ncz;
    float lx, ly, lz, hx, hy, hz, cwx, cwy, cwz, ilx, ily, ilz;
    float cut2;   // FP32 test radius squared
    float cutr;   // radius of the cell-range enumeration (test radius + rounding margin)
    float qzk;    // 2^32 / lz: the z key of the half-list test
};

// z key of a site for the half-list test: z in units of lz / 2^32, kept mod
// 2^32, so the wrap-around difference of two keys is the minimum-image dz
// (to a few 1e-6 A) and exactly minus the difference taken the other way
__device__ __forceinline__ int zkey(float z, float qzk) { return (int)(unsigned long long)__float2ll_rn(z * qzk); }

__device__ __forceinline__ float wave_minf(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// coordinate x in the minimum-image frame of ref
__device__ __forceinline__ float frame1(float x, float ref, float l, float il) {
    float d = x - ref;
    return ref + (d - l * rintf(d * il));
}

struct AxisRange {
    int a, b;   // cell range in the frame's (unwrapped) cell numbering
    bool mi;    // covers the whole ring: visit every cell once, minimum image
};

__device__ __forceinline__ AxisRange axis_range(float mn, float mx, float r, float h, float cw, int nc) {
    int a = (int)floorf((mn - r + h) / cw);
    int b = (int)floorf((mx + r + h) / cw);
    if (b - a + 1 >= nc || a < -nc || b >= 2 * nc) return {0, nc - 1, true};
    return {a, b, false};
}

// the sorted-site ranges (at most two, three with the whole ring) of
// candidate row r; fills lo/len/meta and their count nr
__device__ void row_ranges(int r, const AxisRange &ay, const AxisRange &az, float xmn, float xmx, float ymn,
                          float ymx, float zmn, float zmx, const GbGeom &g, const int *__restrict__ cell_start,
                          const int *__restrict__ cell_cnt, int *lo, int *len, int *meta, int *nr) {
    int nyr = ay.b - ay.a + 1;
    int cyu = ay.a + r % nyr, czu = az.a + r / nyr;
    float gy = 0.0f, gz = 0.0f;
    if (!ay.mi) {
        float s0 = -g.hy + cyu * g.cwy;
        gy = fmaxf(0.0f, fmaxf(s0 - ymx, ymn - (s0 + g.cwy)));
    }
    if (!az.mi) {
        float s0 = -g.hz + czu * g.cwz;
        gz = fmaxf(0.0f, fmaxf(s0 - zmx, zmn - (s0 + g.cwz)));
    }
    float g2 = gy * gy + gz * gz;
    *nr = 0;
    if (g2 > g.cutr * g.cutr) return;
    float xr = sqrtf(g.cutr * g.cutr - g2);
    AxisRange ax = axis_range(xmn, xmx, xr, g.hx, g.cwx, g.ncx);
    int cyw = cyu - floordiv(cyu, g.ncy) * g.ncy, czw = czu - floordiv(czu, g.ncz) * g.ncz;
    int sy = ay.mi ? 0 : floordiv(cyu, g.ncy), sz = az.mi ? 0 : floordiv(czu, g.ncz);
    int rowc = g.ncx * (cyw + g.ncy * czw);
    bool mi = ay.mi || az.mi || ax.mi;
    int p0 = ax.a;
    while (p0 <= ax.b && *nr < 3) {
        int sx = ax.mi ? 0 : floordiv(p0, g.ncx);
        int w0 = p0 - sx * g.ncx;
        int p1 = ax.mi ? ax.b : min(ax.b, (sx + 1) * g.ncx - 1);
        int w1 = p1 - sx * g.ncx;
        int jl = cell_start[rowc + w0];
        int jh = cell_start[rowc + w1] + cell_cnt[rowc + w1];
        if (jh > jl) {
            lo[*nr] = jl;
            len[*nr] = jh - jl;
            meta[*nr] = ((sx + 1) << 7) | ((sy + 1) << 9) | ((sz + 1) << 11) | (mi ? GB_MI : 0);
            (*nr)++;
        }
        p0 = p1 + 1;
    }
}

// Half lists (HalfWin.half != 0): a candidate j is kept in row i only if it
// lies above i: the wrap-around difference of the z keys (zkey: z in units
// of lz / 2^32, mod 2^32; its minimum-image dz to a few 1e-6 A) is > 0, or
// a tie (0 or -2^31) and j > i.  The key difference of (j, i) is exactly minus that of
// (i, j), so every pair is listed once, and each row holds the upper half of
// its sphere (rows of even length, unlike a split by sorted index, which
// follows cell planes).
// Every kept partner sets bit jrel = (j - sb) + pre of an LDS bitmap (sb:
// the group's first site, j - sb taken cyclically into (-n/2, n/2]; pre
// covers the partners of the same cell plane below the group); after the
// rows are complete the bitmap's prefix counts give each distinct partner
// its window slot (ascending cyclic order from sb - pre), the window
// win[g][slot] = j is written, and every entry's j is replaced by its slot.
struct HalfWin {
    int half;        // build a half list with windows
    int span;        // bitmap bits (multiple of 64)
    int pre;         // bitmap bit of the group's first site
    int wcap;        // window capacity
    int *win, *wlen;
    int wlim;        // window length above which an overflow is reported (<= wcap)
    int rlim;        // row length above which an overflow is reported (<= cap)
};

__device__ __forceinline__ void build_grp_body(
    int grp, int n, int cap, const GbGeom &g, const float4 *__restrict__ posfb, const double4 *__restrict__ pos4b,
    const double4 *__restrict__ vel4b, const double4 *__restrict__ f4b, const int *__restrict__ typeb,
    const int *__restrict__ origb, double4 *pos4, double4 *vel4, double4 *f4, int *type, int *orig, double4 *anchor,
    const double4 *__restrict__ fsb, const double4 *__restrict__ flb, double4 *fs, double4 *fl,
    const int *__restrict__ cell_cnt, const int *__restrict__ cell_start, const int *__restrict__ ex_off,
    const int *__restrict__ ex_part, const int *__restrict__ ex_flag, int ex_span, uint32_t *nbl, int *glen,
    int *gcnt, int *status, int conditional, int copy_state, const HalfWin &hw, int4 *posq = nullptr,
    double qkx = 0.0, double qky = 0.0, double qkz = 0.0) {
    if (rebuild_off(status, conditional)) return;
    // partner bitmap (span bits) and its per-word prefix counts
    extern __shared__ unsigned long long hbm[];
    unsigned int *hpre = (unsigned int *)(hbm + (hw.half ? hw.span / 64 : 0));
    const int nword = hw.half ? hw.span / 64 : 0;
    for (int w = threadIdx.x; w < nword; w += blockDim.x) hbm[w] = 0ull;
    __shared__ int srow[64];
    __shared__ int smax;
    // per wave: ring of candidates that passed the box filter (frame
    // position, original index | type bits; sorted index)
    __shared__ float4 sq[GB_WAVES][GB_QCAP];
    __shared__ int sqj[GB_WAVES][GB_QCAP];
    __shared__ int sqz[GB_WAVES][GB_QCAP];   // z keys of the candidates (half lists)
    // exclusion partners of the group's sites (original index) and their
    // class; sxn = count, or -1 where the site has more than GB_EXMAX
    __shared__ int sxo[64][GB_EXMAX];
    __shared__ int sxc[64][GB_EXMAX];
    __shared__ int sxn[64];
    // where every exclusion partner lies within GB_XM_SPAN of its site in the
    // original numbering: the class of partner oj of site oi is 2 bits at
    // 2 * (oj - oi + GB_XM_SPAN) of the site's mask (0 where not listed)
    __shared__ unsigned long long sxm[64];
    const bool xmask = ex_span <= GB_XM_SPAN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sb = grp * 64;
    if (threadIdx.x < 64) {
        int s = sb + threadIdx.x, cnt = 0;
        unsigned long long cm = 0ull;
        if (s < n) {
            int oi = origb[s];
            int x0 = ex_off[oi], x1 = ex_off[oi + 1];
            if (xmask) {
                // later entries of the same partner replace earlier ones
                for (int q = x0; q < x1; q++) {
                    const int f = ex_flag[q];
                    const int sh = 2 * (ex_part[q] - oi + GB_XM_SPAN);
                    cm = (cm & ~(3ull << sh)) | ((unsigned long long)(f == 4 ? 2 : (f ? 1 : 0)) << sh);
                }
            } else {
                cnt = x1 - x0 > GB_EXMAX ? -1 : x1 - x0;
                for (int q = 0; q < cnt; q++) {
                    int f = ex_flag[x0 + q];
                    sxo[threadIdx.x][q] = ex_part[x0 + q];
                    sxc[threadIdx.x][q] = f == 4 ? 2 : (f ? 1 : 0);
                }
            }
        }
        sxn[threadIdx.x] = cnt;
        sxm[threadIdx.x] = cm;
    }
    // the new sorted order becomes the current state
    for (int s = sb + (int)threadIdx.x; copy_state && s < min(sb + 64, n); s += 64 * GB_WAVES) {
        double4 pi = pos4b[s];
        pos4[s] = pi;
        if (posq) posq[s] = ttc_posq(pi, qkx, qky, qkz);
        vel4[s] = vel4b[s];
        f4[s] = f4b[s];
        fs[s] = fsb[s];
        fl[s] = flb[s];
        type[s] = typeb[s];
        orig[s] = origb[s];
        anchor[s] = pi;
    }
    __syncthreads();

    const int row0 = __builtin_amdgcn_readfirstlane(wave) * GB_ROWS;   // this wave's first row in the group
    const int wb = sb + row0;          // and its sorted index
    const int nrw = max(0, min(GB_ROWS, n - wb));
    // lane r < GB_ROWS: the exclusion class mask of row r
    const unsigned long long rcm = sxm[row0 + min(lane, GB_ROWS - 1)];
    int rkl = 0;   // entry count of row lane (lanes below GB_ROWS)
    uint32_t *grow = nbl + (size_t)grp * cap * 64;
    if (nrw > 0) {
        // the frame: minimum image around the wave's first row
        const float4 pref = posfb[wb];
        const float4 pw = posfb[wb + min(lane, nrw - 1)];
        // lane r < nrw holds row r: frame coordinates (ux, uy, uz) and
        // original index (pwo), read wave-uniform with readlane
        const float ux = frame1(pw.x, pref.x, g.lx, g.ilx), uy = frame1(pw.y, pref.y, g.ly, g.ily),
                    uz = frame1(pw.z, pref.z, g.lz, g.ilz);
        const int pwo = __float_as_int(pw.w) & (int)TTC_JMASK;
        const int rowzk = zkey(pw.z, g.qzk);   // z key of row lane
        const float xmn = wave_minf(ux), xmx = wave_maxf(ux);
        const float ymn = wave_minf(uy), ymx = wave_maxf(uy);
        const float zmn = wave_minf(uz), zmx = wave_maxf(uz);
        const AxisRange ay = axis_range(ymn, ymx, g.cutr, g.hy, g.cwy, g.ncy);
        // half lists: a partner lies above some row (minimum-image dz >= 0),
        // so nothing below the rows' lowest z (less a rounding margin) is a
        // candidate; that removes about half of the candidate tests
        const float zlo = zmn - GB_ZMARGIN;
        const AxisRange az = axis_range(hw.half ? zlo + g.cutr : zmn, zmx, g.cutr, g.hz, g.cwz, g.ncz);
        // whole-ring ranges (small boxes): candidates keep their own image
        // and every pair takes the minimum image; no box filter then
        const bool wmi = ay.mi || az.mi || axis_range(xmn, xmx, g.cutr, g.hx, g.cwx, g.ncx).mi;
        const float bcx = 0.5f * (xmn + xmx), bcy = 0.5f * (ymn + ymx), bcz = 0.5f * (zmn + zmx);
        const float bhx = 0.5f * (xmx - xmn), bhy = 0.5f * (ymx - ymn), bhz = 0.5f * (zmx - zmn);
        const float cutr2 = g.cutr * g.cutr;   // box filter radius: the test radius plus a margin
        float4 *qp = sq[row0 / GB_ROWS];
        int *qj = sqj[row0 / GB_ROWS];
        int *qz = sqz[row0 / GB_ROWS];
        int qh = 0, qn = 0;   // ring head and length (wave-uniform)

        // test every row against the cnt (<= 64) candidates at the head of the ring.
        // Lane r holds row r's coordinates and entry count, read wave-uniform;
        // the rows' ballots collect the candidates some row keeps, whose
        // bitmap bits are set once after the rows.
        auto consume = [&](int cnt, auto MI, auto HALF) {
            const int slot = (qh + lane) & (GB_QCAP - 1);
            const float4 pj = qp[slot];
            const int j = qj[slot];
            const int zj = decltype(HALF)::value ? qz[slot] : 0;
            const bool ok = lane < cnt;
            const int wj = __float_as_int(pj.w);
            const int oj = wj & (int)TTC_JMASK;                                    // original index
            const uint32_t ej = (uint32_t)j | ((uint32_t)wj & ~TTC_JMASK);         // sorted index | type bits
            unsigned long long kmask = 0ull;   // candidates some row keeps (half lists)
#pragma unroll 2
            for (int rr = 0; rr < nrw; rr++) {
                const float rwx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ux), rr));
                const float rwy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(uy), rr));
                const float rwz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(uz), rr));
                float dx = rwx - pj.x, dy = rwy - pj.y, dz = rwz - pj.z;
                if (decltype(MI)::value) {
                    dx -= g.lx * rintf(dx * g.ilx);
                    dy -= g.ly * rintf(dy * g.ily);
                    dz -= g.lz * rintf(dz * g.ilz);
                }
                bool hit = ok & (dx * dx + dy * dy + dz * dz < g.cut2);
                if (decltype(HALF)::value) {
                    // the upper half: key difference > 0; ties (0, and -2^31,
                    // its own negative) by index (a site is not above itself)
                    const int dq = zj - __builtin_amdgcn_readlane(rowzk, rr);
                    hit = hit & ((dq > 0) | (((unsigned)dq << 1 == 0u) & (j > wb + rr)));
                } else {
                    hit = hit & (j != wb + rr);
                }
                const unsigned long long b = __builtin_amdgcn_ballot_w64(hit);
                kmask |= b;
                if (b != 0ull) {
                    const int base = __builtin_amdgcn_readlane(rkl, rr);
                    const int pos = base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
                    if (lane == rr) rkl = base + (int)__popcll(b);
                    // exclusion class: only pairs close in the original numbering
                    // can be excluded or 1-4 pairs (max |a-b| over the exclusion list)
                    uint32_t cls = 0;
                    const int oi = __builtin_amdgcn_readlane(pwo, rr);
                    const bool nearo = hit & ((unsigned)(oj - oi + ex_span) <= (unsigned)(2 * ex_span));
                    if (__builtin_amdgcn_ballot_w64(nearo) != 0ull) {
                        if (xmask) {
                            const uint32_t mlo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rcm, rr);
                            const uint32_t mhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(rcm >> 32), rr);
                            const unsigned long long cm = ((unsigned long long)mhi << 32) | mlo;
                            const unsigned d = (unsigned)(oj - oi + GB_XM_SPAN);
                            cls = nearo ? (uint32_t)(cm >> (2u * d)) & 3u : 0u;
                        } else if (nearo) {
                            const int rw = row0 + rr;
                            const int xc = sxn[rw];
                            for (int q2 = 0; q2 < xc; q2++)
                                if (sxo[rw][q2] == oj) cls = (uint32_t)sxc[rw][q2];
                            if (xc < 0) {
                                for (int q2 = ex_off[oi]; q2 < ex_off[oi + 1]; q2++)
                                    if (ex_part[q2] == oj) {
                                        const int f = ex_flag[q2];
                                        cls = f == 4 ? 2u : (f ? 1u : 0u);
                                    }
                            }
                        }
                    }
                    // ttc_nbl_idx in 32-bit arithmetic (a group's block is cap*64 < 2^31 words)
                    if (hit && pos < cap)
                        grow[(uint32_t)(((pos >> 4) << 10) + (row0 + rr) * 16 + (pos & 15))] = ej | (cls << 30);
                }
            }
            if (decltype(HALF)::value && ((kmask >> lane) & 1ull)) {
                // partner's bitmap bit (cyclic offset from the group's first site)
                int djs = j - sb;
                if (2 * djs > n) djs -= n;
                else if (2 * djs <= -n) djs += n;
                const int jrel = djs + hw.pre;
                if (jrel < 0) {
                    atomicMax(&status[ST_PREMAX], -djs);
                    atomicOr(&status[ST_OVERFLOW], 8);
                } else if (jrel >= hw.span) {
                    atomicMax(&status[ST_SPANMAX], jrel + 1);
                    atomicOr(&status[ST_OVERFLOW], 8);
                } else {
                    atomicOr(&hbm[jrel >> 6], 1ull << (jrel & 63));
                }
            }
            qh = (qh + cnt) & (GB_QCAP - 1);
            qn -= cnt;
        };
        auto drain = [&](int cnt) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // ring writes before the reads
            if (hw.half) {
                if (wmi) consume(cnt, std::true_type{}, std::true_type{});
                else consume(cnt, std::false_type{}, std::true_type{});
            } else {
                if (wmi) consume(cnt, std::true_type{}, std::false_type{});
                else consume(cnt, std::false_type{}, std::false_type{});
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // reads before the next writes
        };

        const int nrows = (ay.b - ay.a + 1) * (az.b - az.a + 1);
        for (int r0 = 0; r0 < nrows; r0 += 64) {
            const int r = r0 + lane;
            int lo[3], len[3], meta[3], nr = 0;
            if (r < nrows) row_ranges(r, ay, az, xmn, xmx, ymn, ymx, zmn, zmx, g, cell_start, cell_cnt, lo, len, meta, &nr);
            unsigned long long act = __ballot(nr > 0);
            while (act != 0ull) {
                const int src = __builtin_ctzll(act);
                act &= act - 1ull;
                const int nq = __builtin_amdgcn_readlane(nr, src);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    if (q >= nq) break;
                    const int j0 = __builtin_amdgcn_readlane(lo[q], src);
                    const int jn = __builtin_amdgcn_readlane(len[q], src);
                    const int mt = __builtin_amdgcn_readlane(meta[q], src);
                    // shift into the frame (none for whole-ring ranges or whole-ring waves)
                    const bool shift = !wmi && !(mt & GB_MI);
                    const float shx = shift ? (float)(((mt >> 7) & 3) - 1) * g.lx : 0.0f;
                    const float shy = shift ? (float)(((mt >> 9) & 3) - 1) * g.ly : 0.0f;
                    const float shz = shift ? (float)(((mt >> 11) & 3) - 1) * g.lz : 0.0f;
                    for (int c0 = 0; c0 < jn; c0 += 64) {
                        const bool ok = c0 + lane < jn;
                        const int j = j0 + min(c0 + lane, jn - 1);   // every lane loads inside the range
                        const float4 pj = posfb[j];
                        const float px = pj.x + shx, py = pj.y + shy, pz = pj.z + shz;
                        // keep candidates within the filter radius of the rows' box
                        const float gx = fmaxf(fabsf(px - bcx) - bhx, 0.0f);
                        const float gy = fmaxf(fabsf(py - bcy) - bhy, 0.0f);
                        const float gz = fmaxf(fabsf(pz - bcz) - bhz, 0.0f);
                        const bool keep = ok & (wmi | ((gx * gx + gy * gy + gz * gz < cutr2) &
                                                       (!hw.half | (pz >= zlo))));
                        const unsigned long long m = __builtin_amdgcn_ballot_w64(keep);
                        if (keep) {
                            const int at = (qh + qn + (int)__builtin_amdgcn_mbcnt_hi(
                                                         (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))) &
                                           (GB_QCAP - 1);
                            qp[at] = make_float4(px, py, pz, pj.w);
                            qj[at] = j;
                            qz[at] = zkey(pj.z, g.qzk);
                        }
                        qn += (int)__popcll(m);
                        if (qn >= 64) drain(64);
                    }
                }
            }
        }
        if (qn > 0) drain(qn);
    }

    // row length of the group = longest row; pad the shorter ones
    if (lane < GB_ROWS) srow[row0 + lane] = rkl;
    __syncthreads();
    if (wave == 0) {
        int tot = sb + lane < n ? srow[lane] : 0;
        int full = tot, sum = tot;
        for (int o = 32; o > 0; o >>= 1) {
            full = max(full, __shfl_xor(full, o, 64));
            sum += __shfl_xor(sum, o, 64);
        }
        if (lane == 0) {
            smax = full;
            gcnt[grp] = sum;
            glen[grp] = min(full, cap);
            if (full > hw.rlim) atomicOr(&status[ST_OVERFLOW], 1);
            atomicMax(&status[ST_MAXROW], full);
        }
    }
    __syncthreads();
    int m = min((smax + 15) & ~15, cap);   // whole 16-entry tiles
    for (int r = row0; r < row0 + GB_ROWS; r++) {
        int k0 = min(srow[r], cap);
        for (int k = k0 + lane; k < m; k += 64) grow[ttc_nbl_idx(k, r)] = TTC_PAD;
    }
    if (!hw.half) return;

    // window: prefix counts of the bitmap words (each thread a run of
    // consecutive words, then a scan over the threads), slots in ascending
    // jrel, win[g][slot] = j
    __shared__ int wtot[GB_WAVES];
    const int per = (nword + blockDim.x - 1) / blockDim.x;
    const int w0 = threadIdx.x * per, w1 = min(nword, w0 + per);
    int cnt = 0;
    for (int w = w0; w < w1; w++) cnt += __popcll(hbm[w]);
    // exclusive prefix over the threads: inclusive scan in each wave, then
    // the totals of the waves before
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    int run = inc - cnt, total = 0;
#pragma unroll
    for (int q = 0; q < GB_WAVES; q++) {
        if (q < wave) run += wtot[q];
        total += wtot[q];
    }
    if (threadIdx.x == 0) {
        hw.wlen[grp] = min(total, hw.wcap);
        atomicMax(&status[ST_WMAX], total);
        if (total > hw.wlim) atomicOr(&status[ST_OVERFLOW], 8);
    }
    for (int w = w0; w < w1; w++) {
        hpre[w] = (unsigned int)run;
        run += __popcll(hbm[w]);
    }
    __syncthreads();
    // win[g][slot] = j: one word per wave at a time, lane b its bit b
    int *gwin = hw.win + (size_t)grp * hw.wcap;
    for (int w = wave; w < nword; w += GB_WAVES) {
        const unsigned long long bits = hbm[w];
        if (!((bits >> lane) & 1ull)) continue;
        const int slot = (int)hpre[w] + __popcll(bits & ((1ull << lane) - 1ull));
        if (slot < hw.wcap) {
            int j = sb - hw.pre + w * 64 + lane;
            j = j < 0 ? j + n : (j >= n ? j - n : j);
            gwin[slot] = j;
        }
    }
    // entries: j -> slot (a slot past the capacity is clamped; the overflow
    // flag makes the host grow the window and build again).  The wave's
    // rows are read 64 entries at a time each, all GB_ROWS loads in flight
    // before any is used.
    int lens[GB_ROWS];
    int lmax = 0;
#pragma unroll
    for (int rr = 0; rr < GB_ROWS; rr++) {
        lens[rr] = sb + row0 + rr < n ? min(srow[row0 + rr], cap) : 0;
        lmax = max(lmax, lens[rr]);
    }
    for (int k = lane; k - lane < lmax; k += 64) {
        uint32_t ev[GB_ROWS];
#pragma unroll
        for (int rr = 0; rr < GB_ROWS; rr++)
            ev[rr] = k < lens[rr] ? grow[(uint32_t)(((k >> 4) << 10) + (row0 + rr) * 16 + (k & 15))] : 0u;
#pragma unroll
        for (int rr = 0; rr < GB_ROWS; rr++) {
            if (k >= lens[rr]) continue;
            const uint32_t e = ev[rr];
            int djs = (int)(e & TTC_JMASK) - sb;
            if (2 * djs > n) djs -= n;
            else if (2 * djs <= -n) djs += n;
            const int jrel = djs + hw.pre;
            int slot = 0;
            if (jrel >= 0 && jrel < hw.span) {
                const unsigned long long wbits = hbm[jrel >> 6];
                slot = (int)hpre[jrel >> 6] + __popcll(wbits & ((1ull << (jrel & 63)) - 1ull));
            }
            slot = min(slot, hw.wcap - 1);
            grow[(uint32_t)(((k >> 4) << 10) + (row0 + rr) * 16 + (k & 15))] = (e & ~TTC_JMASK) | (uint32_t)slot;
        }
    }
}

struct BuildGrpPack {
    int n, cap;
    GbGeom g;
    const float4 *posfb;
    const double4 *pos4b, *vel4b, *f4b;
    const int *typeb, *origb;
    double4 *pos4, *vel4, *f4;
    int *type, *orig;
    double4 *anchor;
    const double4 *fsb, *flb;
    double4 *fs, *fl;
    const int *cell_cnt, *cell_start, *ex_off, *ex_part, *ex_flag;
    int ex_span;
    uint32_t *nbl;
    int *glen, *gcnt, *status;
    int conditional;
    int copy_state;
    HalfWin hw;
    int4 *posq;              // fixed-point coordinates of the new order (MD-loop builds)
    double qkx, qky, qkz;
};

__device__ __forceinline__ void build_grp_pack(int grp, const BuildGrpPack &a) {
    build_grp_body(grp, a.n, a.cap, a.g, a.posfb, a.pos4b, a.vel4b, a.f4b, a.typeb, a.origb, a.pos4, a.vel4, a.f4,
                   a.type, a.orig, a.anchor, a.fsb, a.flb, a.fs, a.fl, a.cell_cnt, a.cell_start, a.ex_off, a.ex_part,
                   a.ex_flag, a.ex_span, a.nbl, a.glen, a.gcnt, a.status, a.conditional, a.copy_state, a.hw,
                   a.posq, a.qkx, a.qky, a.qkz);
}
// at most 72 VGPRs: 7 workgroups per CU, so that the groups of about 1800
// x 64 sites (115k) run as one round on 256 CUs
__global__ void __launch_bounds__(64 * GB_WAVES, 7) k_build_grp(BuildGrpPack a) { build_grp_pack(blockIdx.x, a); }
__global__ void __launch_bounds__(64 * GB_WAVES, 7) kb_build_grp(const BuildGrpPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const BuildGrpPack a = pk[m.rep];
    build_grp_pack(m.blk, a);
}

// nblist.c:607-647 on every site; sets the trigger flag.  Also the
// device-side nl_check(lb,ub) of the C ABI (original numbering range).
__global__ void k_nlcheck(int n, const double4 *pos4, const double4 *anchor, const int *orig, int lb,
                          int ub, double lx, double ly, double lz, double rcut, int *status) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int o = orig[s];
    if (o < lb || o >= ub) return;
    double4 p = pos4[s], a = anchor[s];
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double dx = fabs(p.x - a.x);
    if (dx > hx) dx -= lx;
    bool f = dx > rcut;
    double dy = fabs(p.y - a.y);
    if (dy > hy) dy -= ly;
    f = f || dy > rcut;
    double dz = fabs(p.z - a.z);
    if (dz > hz) dz -= lz;
    f = f || dz > rcut;
    f = f || (ttc_r2_exact(dx, dy, dz) > rcut * rcut);
    if (f) atomicOr(&status[ST_NLFLAG], 1);
}

// Reference-layout export: row orig[s] keeps entries with orig[j] > orig[s].
__global__ void k_export(int n, int cap, const uint32_t *nbl, const int *glen, const int *orig,
                         const int *ex_off, const int *ex_part, const int *ex_flag, int max_nb,
                         int *count, int *list, int *xflag, int *status) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int lane = s & 63, grp = s >> 6;
    int a = orig[s];
    int len = glen[grp];
    int c = 0;
    const uint32_t *row = nbl + (size_t)grp * cap * 64;
    for (int k = 0; k < len; k++) {
        uint32_t e = row[ttc_nbl_idx(k, lane)];
        if (e == TTC_PAD) continue;
        int b = orig[e & TTC_JMASK];
        if (b <= a) continue;
        int f = 0;
        for (int q = ex_off[a]; q < ex_off[a + 1]; q++)
            if (ex_part[q] == b) f = ex_flag[q];
        if (c < max_nb) {
            list[(size_t)a * max_nb + c] = b;
            xflag[(size_t)a * max_nb + c] = f;
        }
        c++;
    }
    if (c > max_nb) atomicOr(&status[ST_OVERFLOW], 4);
    count[a] = c;
}

CellGeom geom(const ttc_ctx *c) {
    CellGeom g;
    g.ncx = c->ncx; g.ncy = c->ncy; g.ncz = c->ncz;
    g.lx = c->p.lx; g.ly = c->p.ly; g.lz = c->p.lz;
    g.hx = 0.5 * c->p.lx; g.hy = 0.5 * c->p.ly; g.hz = 0.5 * c->p.lz;
    return g;
}

}  // namespace

double ttc_list_cutoff(const ttc_ctx *c) {
    return c->p.ewald_lists ? std::min(c->p.rclist, c->p.rclist_ew) : c->p.rclist;
}

double ttc_compute_cutoff(const ttc_ctx *c) {
    double pruned = c->p.rc + 0.5 * (c->p.rclist - c->p.rc) + 1e-6;
    return std::min(ttc_list_cutoff(c), pruned);
}

// capacity of one cell's member slots
int ttc_cellcap(const ttc_ctx *c) {
    double vcell = (c->p.lx / c->ncx) * (c->p.ly / c->ncy) * (c->p.lz / c->ncz);
    double avg = c->n * vcell / (c->p.lx * c->p.ly * c->p.lz);
    int cap = (int)(avg * 3.0) + 64;
    return cap;
}

// The per-component early rejects use sim.rclist whichever reference list
// is built (nblist.c:116-127 test against rcut = sim.rclist also while
// filling nlist_ew), so a wider nlist_ew keeps only pairs with every
// |d| <= rclist; lists built inside MD loops reject at their own cutoff.
// margin of the MD-loop (FP32) test, A: far above FP32 rounding of distances
// in boxes up to 1000 A, negligible in list length
#define FAST_MARGIN 0.01

static void launch_build(ttc_ctx *c, bool conditional, double cutoff, uint32_t *nbl, int *glen,
                         int cap, bool copy_state, double component_cut) {
    CellGeom g = geom(c);
    k_build<<<c->G, 64 * BUILD_WAVES, 0, c->stream>>>(
        c->n, cap, g, component_cut, cutoff * cutoff, c->pos4b, c->vel4b, c->f4b, c->typeb, c->origb, c->pos4,
        c->vel4, c->f4, c->type, c->orig, c->anchor, c->inv, c->fsb, c->flb, c->fs, c->fl, c->cell_cnt,
        c->cell_start, c->ex_off, c->ex_part, c->ex_flag, c->ex_span, nbl, glen, c->gcnt, c->status,
        conditional ? 1 : 0, copy_state ? 1 : 0);
}

// the pruned list of MD loops (cutoff ttc_compute_cutoff), FP32 test with
// FAST_MARGIN, group-cooperative build
static BuildGrpPack build_grp_pack_of(ttc_ctx *c, bool half = false, int cond = 1, int copy = 1) {
    GbGeom g;
    g.ncx = c->ncx; g.ncy = c->ncy; g.ncz = c->ncz;
    g.lx = (float)c->p.lx; g.ly = (float)c->p.ly; g.lz = (float)c->p.lz;
    g.hx = 0.5f * g.lx; g.hy = 0.5f * g.ly; g.hz = 0.5f * g.lz;
    g.cwx = (float)(c->p.lx / c->ncx); g.cwy = (float)(c->p.ly / c->ncy); g.cwz = (float)(c->p.lz / c->ncz);
    g.ilx = (float)(1.0 / c->p.lx); g.ily = (float)(1.0 / c->p.ly); g.ilz = (float)(1.0 / c->p.lz);
    double cf = ttc_compute_cutoff(c) + FAST_MARGIN;
    g.cut2 = (float)(cf * cf);
    g.cutr = (float)(cf + FAST_MARGIN);
    g.qzk = (float)(4294967296.0 / c->p.lz);
    return BuildGrpPack{c->n, c->cap, g, c->posfb, c->pos4b, c->vel4b, c->f4b, c->typeb, c->origb, c->pos4, c->vel4,
                        c->f4, c->type, c->orig, c->anchor, c->fsb, c->flb, c->fs, c->fl, c->cell_cnt, c->cell_start,
                        c->ex_off, c->ex_part, c->ex_flag, c->ex_span, c->nbl, c->glen, c->gcnt, c->status, cond,
                        copy, HalfWin{half ? 1 : 0, c->span, c->span_pre, c->wcap, c->win, c->wlen,
                                      c->win_limit > 0 ? std::min(c->win_limit, c->wcap) : c->wcap,
                                      c->row_limit > 0 ? std::min(c->row_limit, c->cap) : c->cap},
                        c->posq, 4294967296.0 / c->p.lx, 4294967296.0 / c->p.ly, 4294967296.0 / c->p.lz};
}

static size_t build_grp_shm(const ttc_ctx *c, bool half) {
    return half ? (size_t)(c->span / 64) * (sizeof(unsigned long long) + sizeof(unsigned int)) : 0;
}

static void launch_build_grp(ttc_ctx *c, bool half, int cond, int copy) {
    k_build_grp<<<c->G, 64 * GB_WAVES, build_grp_shm(c, half), c->stream>>>(build_grp_pack_of(c, half, cond, copy));
}

static CellAssignPack cell_assign_pack(ttc_ctx *c, int cond) {
    return CellAssignPack{c->n, (c->n + 255) / 256, c->pos4, geom(c), ttc_cellcap(c), c->ncell, c->cell_acc,
                          c->cell_cnt, c->cell_start, c->cell_of, c->cell_members, c->status, cond};
}

static ReorderPack reorder_pack(ttc_ctx *c, int cond) {
    return ReorderPack{c->n, ttc_cellcap(c), c->cell_of, c->cell_cnt, c->cell_start, c->cell_members, c->orig,
                       c->pos4, c->vel4, c->f4, c->type, c->pos4b, c->vel4b, c->f4b, c->typeb, c->origb, c->fs,
                       c->fl, c->fsb, c->flb, c->posfb, c->inv, c->status, cond};
}

int ttc_launch_rebuild(ttc_ctx *c, bool conditional) {
    int blocks = (c->n + 255) / 256;
    int cond = conditional ? 1 : 0;
    k_cell_assign<<<blocks, 256, 0, c->stream>>>(cell_assign_pack(c, cond));
    k_reorder<<<blocks, 256, 0, c->stream>>>(reorder_pack(c, cond));
    // explicit builds (nblist.c:32) make the reference's list; the
    // conditional builds of MD loops make the pruned one
    if (conditional)
        launch_build_grp(c, c->half_mode, 1, 1);
    else
        launch_build(c, false, ttc_list_cutoff(c), c->nbl, c->glen, c->cap, true, c->p.rclist);
    return TTC_OK;
}

// the builder's dynamic LDS (bitmap) may exceed 64 KiB in very large boxes
int ttc_build_grp_attrs(ttc_ctx *c) {
    int shm = (int)build_grp_shm(c, true);
    TTC_HIP(hipFuncSetAttribute((const void *)k_build_grp, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    return TTC_OK;
}

int ttc_launch_rebuild_md(ttc_ctx *c) {
    int blocks = (c->n + 255) / 256;
    k_cell_assign<<<blocks, 256, 0, c->stream>>>(cell_assign_pack(c, 0));
    k_reorder<<<blocks, 256, 0, c->stream>>>(reorder_pack(c, 0));
    launch_build_grp(c, c->half_mode, 0, 1);
    return TTC_OK;
}

// posfb holds the FP32 positions the current sorted order was made from,
// i.e. the anchors
int ttc_launch_build_md_at_anchors(ttc_ctx *c, bool half) {
    launch_build_grp(c, half, 0, 0);
    return TTC_OK;
}

// Window capacity: the distinct partners of a group are the sites within
// the list radius rp of the group's box above its lowest site, about half of
// rho (V + S rp + pi (a + b + c) rp^2 + 4/3 pi rp^3) for a box a x b x c
// (64 consecutive sorted sites: along x over whole cells of the y-z
// cross-section); 40 % headroom.  Bitmap: a partner lies in the cell plane
// of its row (z, the slowest cell axis), up to one plane of sites before the
// group's first site, or in the two planes above.
void ttc_window_sizes(const ttc_ctx *c, int *wcap, int *span, int *pre) {
    double vol = c->p.lx * c->p.ly * c->p.lz;
    double rho = c->n / vol;
    double cwx = c->p.lx / c->ncx, cwy = c->p.ly / c->ncy, cwz = c->p.lz / c->ncz;
    double rp = ttc_compute_cutoff(c) + 2.0 * FAST_MARGIN;
    double a = std::min(c->p.lx, std::max(cwx, 64.0 / (rho * cwy * cwz))), b = cwy, cc = cwz;
    double vm = a * b * cc + 2.0 * (a * b + b * cc + cc * a) * rp + TTC_PI * (a + b + cc) * rp * rp +
                4.0 / 3.0 * TTC_PI * rp * rp * rp;
    double est = 0.5 * rho * vm + 64.0;
    int w = ((int)(1.4 * est) + 63) & ~63;
    *wcap = std::max(64, std::min(w, ((c->n + 63) / 64) * 64));
    ttc_bitmap_sizes(c, (int)(1.3 * c->n / c->ncz) + 64, (int)(3.3 * c->n / c->ncz) + 128, span, pre);
}

// bitmap of the builder for partners from `below` sites before a group's
// first site to `above` after it, rounded to 64 bits; in boxes of fewer
// than 5 cell planes along z (or when that covers the box) every site
void ttc_bitmap_sizes(const ttc_ctx *c, int below, int above, int *span, int *pre) {
    long long b = ((long long)below + 63) & ~63LL;
    long long s = (b + above + 63) & ~63LL;
    if (c->ncz < 5 || s >= c->n) {
        b = ((long long)c->n / 2 + 64 + 63) & ~63LL;
        s = (2 * b + 63) & ~63LL;
    }
    *pre = (int)b;
    *span = (int)s;
}

// the reference's list (full cutoff) at the anchors, in the current sorted
// order (made from the anchors), without touching the site state
int ttc_launch_build_at_anchors(ttc_ctx *c) {
    TTC_HIP(hipMemcpyAsync(c->pos4b, c->anchor, sizeof(double4) * c->n, hipMemcpyDeviceToDevice, c->stream));
    TTC_HIP(hipMemcpyAsync(c->origb, c->orig, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
    launch_build(c, false, ttc_list_cutoff(c), c->nbl, c->glen, c->cap, false, c->p.rclist);
    return TTC_OK;
}

void ttc_launch_nlcheck(ttc_ctx *c) {
    int blocks = (c->n + 255) / 256;
    k_nlcheck<<<blocks, 256, 0, c->stream>>>(c->n, c->pos4, c->anchor, c->orig, 0, c->n, c->p.lx, c->p.ly,
                                             c->p.lz, (c->p.rclist - c->p.rc) / 4.0, c->status);
}

int ttc_nl_check_range(ttc_ctx *c, int lb, int ub, int *flag) {
    TTC_HIP(hipMemsetAsync(c->status + ST_NLFLAG, 0, sizeof(int), c->stream));
    int blocks = (c->n + 255) / 256;
    k_nlcheck<<<blocks, 256, 0, c->stream>>>(c->n, c->pos4, c->anchor, c->orig, lb, ub, c->p.lx, c->p.ly,
                                             c->p.lz, (c->p.rclist - c->p.rc) / 4.0, c->status);
    TTC_HIP(hipMemcpyAsync(c->h_status, c->status, sizeof(int) * ST_NSTAT, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    *flag = c->h_status[ST_NLFLAG];
    return TTC_OK;
}

int ttc_export_list(ttc_ctx *c, double cutoff, int max_nb, int *count, int *list, int *xflag) {
    if (!c->have_list) { ttc_set_error("ttc_get_nblist: no list built (call ttc_nblist first)"); return TTC_ERR_ARG; }
    uint32_t *nbl = c->nbl;
    int *glen = c->glen;
    int cap = c->cap;
    uint32_t *tmp_nbl = nullptr;
    int *tmp_glen = nullptr;
    if (cutoff != ttc_list_cutoff(c) || c->list_pruned) {
        // Another cutoff (nlist_ew wider than nlist): build a scratch list at
        // the anchor positions, i.e. the positions of the last ttc_nblist, for
        // which the current sorted order and cell ranges were made.  Cells
        // are sized for the wider cutoff at creation.
        double ratio = std::max(1.0, cutoff / ttc_list_cutoff(c));
        cap = ((int)(c->cap * ratio * ratio * ratio) + 127) & ~63;
        TTC_HIP(hipMalloc(&tmp_nbl, (size_t)c->G * cap * 64 * sizeof(uint32_t)));
        TTC_HIP(hipMalloc(&tmp_glen, (size_t)c->G * sizeof(int)));
        TTC_HIP(hipMemcpyAsync(c->pos4b, c->anchor, sizeof(double4) * c->n, hipMemcpyDeviceToDevice, c->stream));
        TTC_HIP(hipMemcpyAsync(c->origb, c->orig, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
        launch_build(c, false, cutoff, tmp_nbl, tmp_glen, cap, false, c->p.rclist);
        nbl = tmp_nbl;
        glen = tmp_glen;
    }
    int *d_count = nullptr, *d_list = nullptr, *d_xf = nullptr;
    size_t tot = (size_t)c->n * max_nb;
    TTC_HIP(hipMalloc(&d_count, sizeof(int) * c->n));
    TTC_HIP(hipMalloc(&d_list, sizeof(int) * tot));
    TTC_HIP(hipMalloc(&d_xf, sizeof(int) * tot));
    TTC_HIP(hipMemsetAsync(c->status + ST_OVERFLOW, 0, sizeof(int), c->stream));
    int blocks = (c->n + 255) / 256;
    k_export<<<blocks, 256, 0, c->stream>>>(c->n, cap, nbl, glen, c->orig, c->ex_off, c->ex_part, c->ex_flag,
                                            max_nb, d_count, d_list, d_xf, c->status);
    TTC_HIP(hipMemcpyAsync(c->h_status, c->status, sizeof(int) * ST_NSTAT, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipMemcpyAsync(count, d_count, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipMemcpyAsync(list, d_list, sizeof(int) * tot, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipMemcpyAsync(xflag, d_xf, sizeof(int) * tot, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    hipFree(d_count); hipFree(d_list); hipFree(d_xf);
    if (tmp_nbl) hipFree(tmp_nbl);
    if (tmp_glen) hipFree(tmp_glen);
    int overflow = c->h_status[ST_OVERFLOW];
    c->h_status[ST_OVERFLOW] = 0;
    if (overflow & 4) {
        ttc_set_error("ttc_get_nblist: a row exceeds max_nb (nblist.c:185-196)");
        return TTC_ERR_LIST_OVERFLOW;
    }
    if (overflow & 3) {
        ttc_set_error("ttc_get_nblist: device list overflow");
        return TTC_ERR_LIST_OVERFLOW;
    }
    // reference row order: ascending b (nblist.c:107, b = a+1 .. ns-1)
    std::vector<std::pair<int, int>> tmp;
    for (int a = 0; a < c->n; a++) {
        int k = count[a];
        tmp.resize(k);
        for (int j = 0; j < k; j++) tmp[j] = {list[(size_t)a * max_nb + j], xflag[(size_t)a * max_nb + j]};
        std::sort(tmp.begin(), tmp.end());
        for (int j = 0; j < k; j++) {
            list[(size_t)a * max_nb + j] = tmp[j].first;
            xflag[(size_t)a * max_nb + j] = tmp[j].second;
        }
    }
    return TTC_OK;
}

// ---- batched stages of the NHC step (ttc_batch.hip): the conditional
// rebuild of every context, st[0..2] = assign (+ scan), reorder, build

int ttc_bstage_rebuild(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<CellAssignPack> pa;
    std::vector<ReorderPack> pr;
    std::vector<BuildGrpPack> pb;
    std::vector<int> nsite, ngrp;
    size_t shm = 0;   // half lists: the largest partner bitmap
    for (ttc_ctx *c : r) {
        pa.push_back(cell_assign_pack(c, 1));
        pr.push_back(reorder_pack(c, 1));
        pb.push_back(build_grp_pack_of(c, c->half_mode, 1, 1));
        nsite.push_back((c->n + 255) / 256);
        ngrp.push_back(c->G);
        shm = std::max(shm, build_grp_shm(c, c->half_mode));
    }
    TTC_HIP(hipFuncSetAttribute((const void *)kb_build_grp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    int e = ttc_stage_make(&st[0], pa, nsite, 0);
    if (!e) e = ttc_stage_make(&st[1], pr, nsite, 0);
    if (!e) e = ttc_stage_make(&st[2], pb, ngrp, shm);
    return e;
}

void ttc_blaunch_rebuild(const TtcStage *st, hipStream_t s) {
    kb_cell_assign<<<st[0].nblk, 256, 0, s>>>((const CellAssignPack *)st[0].pk, st[0].map);
    kb_reorder<<<st[1].nblk, 256, 0, s>>>((const ReorderPack *)st[1].pk, st[1].map);
    kb_build_grp<<<st[2].nblk, 64 * GB_WAVES, st[2].shm, s>>>((const BuildGrpPack *)st[2].pk, st[2].map);
}
