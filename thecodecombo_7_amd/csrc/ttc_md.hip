// ttc_md.hip -- force combination, second half-kick, kinetic energy and the
// deterministic reductions (integrate_vverlet.c:93-102, kinet.c:18-68,
// pme_calc.c:268-277 mean PME force removal, nblist.c:607-647 trigger).
//
// The per-site force of one evaluation is the sum of the pair kernel's
// force, the gather's PME + bond force, minus the mean PME force; the three
// are added as int64 fixed point (2^-32 kJ/mol/A), so the total does not
// depend on the order in which contributions arrive.
#include "ttc_common.h"

namespace {

// mean PME force over all sites from the gather's per-block sums, reduced
// in a fixed order (lane-strided, then a shuffle tree) by wave 0
__device__ __forceinline__ void mean_pme(const double *partials, int maxb, int nbg, int n, double *sh3) {
    if (threadIdx.x < 64) {
        for (int d = 0; d < 3; d++) {
            double t = 0.0;
            for (int b = threadIdx.x; b < nbg; b += 64) t += partials[(size_t)(EV_FSX + d) * maxb + b];
            t = ttc_wave_sum(t);
            if (threadIdx.x == 0) sh3[d] = t / (double)n;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double4 combine(int s, int mask, const double4 *fpair, const double4 *g4,
                                           const double *mean) {
    long long fx = 0, fy = 0, fz = 0;
    if (mask & 1) {
        double4 a = fpair[s];
        fx += ttc_to_fix(a.x); fy += ttc_to_fix(a.y); fz += ttc_to_fix(a.z);
    }
    if (mask & 2) {
        double4 b = g4[s];
        fx += ttc_to_fix(b.x); fy += ttc_to_fix(b.y); fz += ttc_to_fix(b.z);
    }
    if (mask & 4) {
        fx -= ttc_to_fix(mean[0]); fy -= ttc_to_fix(mean[1]); fz -= ttc_to_fix(mean[2]);
    }
    return make_double4((double)fx * TTC_FIX_INV, (double)fy * TTC_FIX_INV, (double)fz * TTC_FIX_INV, 0.0);
}

__global__ void __launch_bounds__(256)
k_combine(int n, int mask, const double4 *fpair, const double4 *g4, double4 *f4, const double *partials,
          int maxb, int nbg) {
    __shared__ double mean[3];
    if (mask & 4) mean_pme(partials, maxb, nbg, n, mean);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    f4[s] = combine(s, mask, fpair, g4, mean);
}

// kinet.c:18-68 with uu = vv: ken = 0.125 m (v+u)^2 * NA*AU*10
__device__ __forceinline__ void ken_block(double m, double4 v, double *partials, int maxb, bool valid) {
    double k[6] = {0, 0, 0, 0, 0, 0};
    if (valid) {
        double sx = v.x + v.x, sy = v.y + v.y, sz = v.z + v.z;
        k[0] = m * sx * sx; k[1] = m * sy * sx; k[2] = m * sy * sy;
        k[3] = m * sz * sx; k[4] = m * sz * sy; k[5] = m * sz * sz;
    }
    __shared__ double red[4][6];
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = 0; q < 6; q++) {
        double t = ttc_wave_sum(k[q]);
        if (lane == 0) red[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        partials[(size_t)(EV_KE0 + threadIdx.x) * maxb + blockIdx.x] = t * (0.125 * (TTC_NA * TTC_AU * 10.0));
    }
}

// integrate_vverlet.c:93-102 (second half kick) on the combined force, then
// the DLIST check of nblist.c:607-647 on the new positions (it decides the
// rebuild at the start of the next step, as nve_md.c:369-370 does).
template <bool ENERGY>
__global__ void __launch_bounds__(256)
k_kick(int n, const double4 *pos4, double4 *vel4, const double4 *fpair, const double4 *g4, double4 *f4,
       const double4 *anchor, double dt, double vscal, const double *pcheck, double *partials, int maxb,
       int nbg, int *status) {
    __shared__ double mean[3];
    mean_pme(partials, maxb, nbg, n, mean);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) status[ST_STEP] += 1;
    double4 v = make_double4(0, 0, 0, 1);
    bool ok = s < n;
    if (ok) {
        double4 f = combine(s, 7, fpair, g4, mean);
        f4[s] = f;
        v = vel4[s];
        double dth = 0.5 * dt;
        v.x = v.x + vscal * dth * v.w * f.x;
        v.y = v.y + vscal * dth * v.w * f.y;
        v.z = v.z + vscal * dth * v.w * f.z;
        vel4[s] = v;
        // displacement trigger
        double4 p = pos4[s], a = anchor[s];
        double lx = pcheck[0], ly = pcheck[1], lz = pcheck[2], rcut = pcheck[3];
        double dx = fabs(p.x - a.x);
        if (dx > 0.5 * lx) dx -= lx;
        double dy = fabs(p.y - a.y);
        if (dy > 0.5 * ly) dy -= ly;
        double dz = fabs(p.z - a.z);
        if (dz > 0.5 * lz) dz -= lz;
        if (dx > rcut || dy > rcut || dz > rcut || ttc_r2_exact(dx, dy, dz) > rcut * rcut)
            atomicOr(&status[ST_NLFLAG], 1);
    }
    if (ENERGY) ken_block(100.0 / v.w, v, partials, maxb, ok);
}

__global__ void __launch_bounds__(256)
k_kinet(int n, const double4 *vel4, double *partials, int maxb) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = s < n;
    double4 v = ok ? vel4[s] : make_double4(0, 0, 0, 1);
    ken_block(100.0 / v.w, v, partials, maxb, ok);
}

struct ReduceArgs {
    int nb[EV_N];
};

// one workgroup: every slot summed over its producer's blocks in a fixed
// order; optionally appends an energy row
// {step, kinet, bond, nbond, nbonds, coulomb, kewald, rewald, iewald}
__global__ void __launch_bounds__(256)
k_reduce(ReduceArgs ra, const double *partials, int maxb, double *ev, double e_self, double fan,
         double *rows, int rows_cap, int *status, int append) {
    __shared__ double red[4];
    for (int slot = 0; slot < EV_N; slot++) {
        double t = 0.0;
        int nb = ra.nb[slot];
        for (int b = threadIdx.x; b < nb; b += 256) t += partials[(size_t)slot * maxb + b];
        t = ttc_wave_sum(t);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
        __syncthreads();
        // the influence-function slots are sums of eterm |S|^2: scale by
        // 0.5 QFACTOR / epsRF (pme_calc.c:95-111)
        if (threadIdx.x == 0)
            ev[slot] = (red[0] + red[1] + red[2] + red[3]) * ((slot >= EV_REC && slot <= EV_KV5) ? fan : 1.0);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ev[EV_KIN] = ev[EV_KE0] + ev[EV_KE2] + ev[EV_KE5];
        if (append) {
            int r = status[ST_NROWS];
            if (r < rows_cap) {
                double erec = ev[EV_REC];
                double ktot = erec - e_self;
                double nbond = ev[EV_LJ] + ev[EV_REAL] - ev[EV_INTRA] + ktot;
                double *row = rows + (size_t)r * ROW_LEN;
                row[0] = (double)status[ST_STEP];
                row[1] = ev[EV_KIN];
                row[2] = ev[EV_BOND];
                row[3] = nbond;
                row[4] = nbond - ev[EV_LJS];
                row[5] = ev[EV_REAL] + ktot - ev[EV_INTRA];
                row[6] = erec;
                row[7] = ev[EV_REAL];
                row[8] = ev[EV_INTRA];
            }
            status[ST_NROWS] = r + 1;
        }
    }
}

}  // namespace

int ttc_kick_blocks(ttc_ctx *c) { return (c->n + 255) / 256; }

void ttc_launch_combine(ttc_ctx *c, int mask) {
    int blocks = (c->n + 255) / 256;
    k_combine<<<blocks, 256, 0, c->stream>>>(c->n, mask, c->fpair, c->g4, c->f4, c->partials, c->nblk_max,
                                             ttc_gather_blocks(c));
}

// device constants for the trigger: {lx, ly, lz, (rclist-rc)/4}
static double *check_consts(ttc_ctx *c) {
    return c->ev + EV_N;   // four doubles stored after the reduced slots
}

void ttc_launch_kick(ttc_ctx *c, double vscal, bool energy) {
    int blocks = ttc_kick_blocks(c);
    if (energy)
        k_kick<true><<<blocks, 256, 0, c->stream>>>(c->n, c->pos4, c->vel4, c->fpair, c->g4, c->f4, c->anchor,
                                                    c->p.dt, vscal, check_consts(c), c->partials, c->nblk_max,
                                                    ttc_gather_blocks(c), c->status);
    else
        k_kick<false><<<blocks, 256, 0, c->stream>>>(c->n, c->pos4, c->vel4, c->fpair, c->g4, c->f4, c->anchor,
                                                     c->p.dt, vscal, check_consts(c), c->partials, c->nblk_max,
                                                     ttc_gather_blocks(c), c->status);
}

void ttc_launch_kinet(ttc_ctx *c) {
    k_kinet<<<ttc_kick_blocks(c), 256, 0, c->stream>>>(c->n, c->vel4, c->partials, c->nblk_max);
}

void ttc_launch_reduce(ttc_ctx *c, bool append_row) {
    ReduceArgs ra;
    for (int i = 0; i < EV_N; i++) ra.nb[i] = 0;
    for (int i = EV_LJ; i <= EV_PV5; i++) ra.nb[i] = ttc_pair_blocks(c);
    for (int i = EV_REC; i <= EV_KV5; i++) ra.nb[i] = ttc_pme_conv_blocks(c);
    for (int i = EV_BOND; i <= EV_BV5; i++) ra.nb[i] = ttc_gather_blocks(c);
    for (int i = EV_KE0; i <= EV_KE5; i++) ra.nb[i] = ttc_kick_blocks(c);
    for (int i = EV_FSX; i <= EV_FSZ; i++) ra.nb[i] = ttc_gather_blocks(c);
    double fan = 0.5 * TTC_QFACTOR / c->p.eps_rf;
    k_reduce<<<1, 256, 0, c->stream>>>(ra, c->partials, c->nblk_max, c->ev, c->e_self, fan, c->rows, c->rows_cap,
                                       c->status, append_row ? 1 : 0);
}
