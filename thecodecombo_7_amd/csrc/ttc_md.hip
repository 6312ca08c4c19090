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

__global__ void __launch_bounds__(256)
k_combine(int n, int mask, const double4 *fpair, const double4 *g4, double4 *f4, const double *partials,
          int maxb, int nbg, long long *fixj) {
    __shared__ double mean[3];
    if (mask & 4) ttc_mean_pme(partials, maxb, nbg, n, mean);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    f4[s] = ttc_combine(s, mask, fpair, g4, mean, -1, fixj, n);
}

// kinet.c:18-68 with uu = vv: ken = 0.125 m (v+u)^2 * NA*AU*10
__device__ __forceinline__ void ken_block(int bid, double m, double4 v, double *partials, int maxb, bool valid) {
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
        partials[(size_t)(EV_KE0 + threadIdx.x) * maxb + bid] = t * (0.125 * (TTC_NA * TTC_AU * 10.0));
    }
}

// integrate_vverlet.c:93-102 (second half kick) on the combined force.  The
// DLIST flag this step's drift set has been consumed by this step's
// conditional rebuild: it is cleared here.
template <bool ENERGY>
__device__ __forceinline__ void kick_body(int bid, int n, double4 *vel4, const double4 *fpair, const double4 *g4,
                                          double4 *f4, double dt, double vscal, double *partials, int maxb, int nbg,
                                          int *status, const int *gorig = nullptr, long long *fixj = nullptr,
                                          const int *ovf = nullptr) {
    // an MD loop whose rebuild overflowed stops here; the host grows the
    // list and redoes the step
    if (ovf && (*ovf & TTC_OVF_REDO)) return;
    __shared__ double mean[3];
    ttc_mean_pme(partials, maxb, nbg, n, mean);
    int s = bid * blockDim.x + threadIdx.x;
    if (s == 0 && status) {
        status[ST_STEP] += 1;
        status[ST_NLFLAG] = 0;
    }
    double4 v = make_double4(0, 0, 0, 1);
    bool ok = s < n;
    if (ok) {
        double4 f = ttc_combine(s, 7, fpair, g4, mean, gorig ? gorig[s] : s, fixj, n);
        f4[s] = f;
        v = vel4[s];
        double dth = 0.5 * dt;
        v.x = v.x + vscal * dth * v.w * f.x;
        v.y = v.y + vscal * dth * v.w * f.y;
        v.z = v.z + vscal * dth * v.w * f.z;
        vel4[s] = v;
    }
    if (ENERGY) ken_block(bid, 100.0 / v.w, v, partials, maxb, ok);
}

struct KickPack {
    int n;
    double4 *vel4;
    const double4 *fpair, *g4;
    double4 *f4;
    double dt, vscal;
    double *partials;
    int maxb, nbg;
    int *status;
    const int *gorig = nullptr;   // g4 in the original numbering: read at gorig[s]
    long long *fixj = nullptr;    // partner forces of a half-list pair kernel
    const int *ovf = nullptr;     // overflow flags of an MD loop (stop when set)
};

template <bool ENERGY>
__global__ void __launch_bounds__(256) k_kick(KickPack a) {
    kick_body<ENERGY>(blockIdx.x, a.n, a.vel4, a.fpair, a.g4, a.f4, a.dt, a.vscal, a.partials, a.maxb, a.nbg, a.status,
                      a.gorig, a.fixj, a.ovf);
}
template <bool ENERGY>
__global__ void __launch_bounds__(256) kb_kick(const KickPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const KickPack a = pk[m.rep];
    kick_body<ENERGY>(m.blk, a.n, a.vel4, a.fpair, a.g4, a.f4, a.dt, a.vscal, a.partials, a.maxb, a.nbg, a.status,
                      a.gorig, a.fixj, a.ovf);
}

__device__ __forceinline__ void kinet_body(int bid, int n, const double4 *vel4, double *partials, int maxb) {
    int s = bid * blockDim.x + threadIdx.x;
    bool ok = s < n;
    double4 v = ok ? vel4[s] : make_double4(0, 0, 0, 1);
    ken_block(bid, 100.0 / v.w, v, partials, maxb, ok);
}

// drift.c:36-44: block sums of m v (m = 100 / vel4.w)
__global__ void __launch_bounds__(256) k_momentum(int n, const double4 *vel4, double *partials, int maxb) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    double p[3] = {0.0, 0.0, 0.0};
    if (s < n) {
        double4 v = vel4[s];
        double m = 100.0 / v.w;
        p[0] = m * v.x; p[1] = m * v.y; p[2] = m * v.z;
    }
    __shared__ double red[4][3];
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = 0; q < 3; q++) {
        double t = ttc_wave_sum(p[q]);
        if (lane == 0) red[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int q = threadIdx.x;
        partials[(size_t)(EV_MOX + q) * maxb + blockIdx.x] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
    }
}

// drift.c:52-72: every block sums the momentum partials in the same order,
// then v_i -= (P / N) / m_i
__global__ void __launch_bounds__(256) k_momentum_remove(int n, double4 *vel4, const double *partials, int maxb,
                                                         int nb) {
    __shared__ double P[3];
    const int d = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (d < 3) {
        double t = 0.0;
        for (int b = lane; b < nb; b += 64) t += partials[(size_t)(EV_MOX + d) * maxb + b];
        t = ttc_wave_sum(t);
        if (lane == 0) P[d] = t / (1.0 * n);
    }
    __syncthreads();
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double4 v = vel4[s];
    double im = 1.0 / (100.0 / v.w);
    v.x -= P[0] * im; v.y -= P[1] * im; v.z -= P[2] * im;
    vel4[s] = v;
}

struct KinetPack {
    int n;
    const double4 *vel4;
    double *partials;
    int maxb;
};

__global__ void __launch_bounds__(256) k_kinet(KinetPack a) { kinet_body(blockIdx.x, a.n, a.vel4, a.partials, a.maxb); }
__global__ void __launch_bounds__(256) kb_kinet(const KinetPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const KinetPack a = pk[m.rep];
    kinet_body(m.blk, a.n, a.vel4, a.partials, a.maxb);
}

struct ReduceArgs {
    int nb[EV_N];
};

// one workgroup: every slot summed over its producer's blocks in a fixed
// order; optionally appends an energy row
// {step, kinet, bond, nbond, nbonds, coulomb, kewald, rewald, iewald}, where
// bond is bond + bend + urey + tors + impr
__device__ __forceinline__ void reduce_body(const ReduceArgs &ra, const double *partials, int maxb, double *ev,
                                            double e_self, double fan, double *rows, int rows_cap, int *status,
                                            int append) {
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
        // no row for a step an overflow stopped (the host's redo appends it)
        if (append && !(status[ST_OVERFLOW] & TTC_OVF_REDO)) {
            int r = status[ST_NROWS];
            if (r < rows_cap) {
                double erec = ev[EV_REC];
                double ktot = erec - e_self;
                double nbond = ev[EV_LJ] + ev[EV_REAL] - ev[EV_INTRA] + ktot;
                double *row = rows + (size_t)r * ROW_LEN;
                row[0] = (double)status[ST_STEP];
                row[1] = ev[EV_KIN];
                // calcvalue.c:44-53: the bonded part of potens
                row[2] = ev[EV_BOND] + ev[EV_EBEND] + ev[EV_EUREY] + ev[EV_ETORS] + ev[EV_EIMPR];
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

struct ReducePack {
    ReduceArgs ra;
    const double *partials;
    int maxb;
    double *ev;
    double e_self, fan;
    double *rows;
    int rows_cap;
    int *status;
    int append;
};

__global__ void __launch_bounds__(256) k_reduce(ReducePack a) {
    reduce_body(a.ra, a.partials, a.maxb, a.ev, a.e_self, a.fan, a.rows, a.rows_cap, a.status, a.append);
}
__global__ void __launch_bounds__(256) kb_reduce(const ReducePack *pk, const TtcBmap *map) {
    const ReducePack &a = pk[map[blockIdx.x].rep];
    reduce_body(a.ra, a.partials, a.maxb, a.ev, a.e_self, a.fan, a.rows, a.rows_cap, a.status, a.append);
}

}  // namespace

int ttc_kick_blocks(ttc_ctx *c) { return (c->n + 255) / 256; }

void ttc_launch_combine(ttc_ctx *c, int mask) {
    int blocks = (c->n + 255) / 256;
    k_combine<<<blocks, 256, 0, c->stream>>>(c->n, mask, c->fpair, c->g4, c->f4, c->partials, c->nblk_max,
                                             ttc_gather_blocks(c), c->fixj);
}

// device constants for the trigger: {lx, ly, lz, (rclist-rc)/4}
static double *check_consts(ttc_ctx *c) {
    return c->ev + EV_N;   // four doubles stored after the reduced slots
}

static KickPack kick_pack(ttc_ctx *c, double vscal) {
    return KickPack{c->n, c->vel4, c->fpair, c->g4, c->f4, c->p.dt, vscal, c->partials, c->nblk_max,
                    ttc_gather_blocks(c), c->status, nullptr, c->fixj, nullptr};
}

void ttc_launch_kick(ttc_ctx *c, double vscal, bool energy) {
    int blocks = ttc_kick_blocks(c);
    if (energy) k_kick<true><<<blocks, 256, 0, c->stream>>>(kick_pack(c, vscal));
    else k_kick<false><<<blocks, 256, 0, c->stream>>>(kick_pack(c, vscal));
}

void ttc_launch_kick_tail(ttc_ctx *c, double vscal, bool energy) {
    int blocks = ttc_kick_blocks(c);
    KickPack a = kick_pack(c, vscal);
    a.status = nullptr;
    a.g4 = c->g4o;
    a.gorig = c->orig;
    a.ovf = c->status + ST_OVERFLOW;
    if (energy) k_kick<true><<<blocks, 256, 0, c->stream>>>(a);
    else k_kick<false><<<blocks, 256, 0, c->stream>>>(a);
}

static KinetPack kinet_pack(ttc_ctx *c) { return KinetPack{c->n, c->vel4, c->partials, c->nblk_max}; }

void ttc_launch_kinet(ttc_ctx *c) {
    k_kinet<<<ttc_kick_blocks(c), 256, 0, c->stream>>>(kinet_pack(c));
}

void ttc_launch_momentum_removal(ttc_ctx *c) {
    int blocks = ttc_kick_blocks(c);
    k_momentum<<<blocks, 256, 0, c->stream>>>(c->n, c->vel4, c->partials, c->nblk_max);
    k_momentum_remove<<<blocks, 256, 0, c->stream>>>(c->n, c->vel4, c->partials, c->nblk_max, blocks);
}

static ReducePack reduce_pack(ttc_ctx *c, bool append_row) {
    ReduceArgs ra;
    for (int i = 0; i < EV_N; i++) ra.nb[i] = 0;
    for (int i = EV_LJ; i <= EV_PV5; i++) ra.nb[i] = ttc_pair_blocks(c);
    for (int i = EV_REC; i <= EV_KV5; i++) ra.nb[i] = ttc_pme_conv_blocks(c);
    for (int i = EV_BOND; i <= EV_BV5; i++) ra.nb[i] = ttc_gather_blocks(c);
    for (int i = EV_KE0; i <= EV_KE5; i++) ra.nb[i] = ttc_kick_blocks(c);
    for (int i = EV_FSX; i <= EV_FSZ; i++) ra.nb[i] = ttc_gather_blocks(c);
    for (int i = EV_PHX; i <= EV_PHZ; i++) ra.nb[i] = ttc_pair_blocks(c);
    for (int i = EV_KHX; i <= EV_BHZ; i++) ra.nb[i] = ttc_gather_blocks(c);
    for (int i = EV_EBEND; i <= EV_AHZ; i++) ra.nb[i] = ttc_angle_blocks(c);
    for (int i = EV_MOX; i <= EV_MOZ; i++) ra.nb[i] = ttc_kick_blocks(c);
    double fan = 0.5 * TTC_QFACTOR / c->p.eps_rf;
    return ReducePack{ra, c->partials, c->nblk_max, c->ev, c->e_self, fan, c->rows, c->rows_cap, c->status,
                      append_row ? 1 : 0};
}

void ttc_launch_reduce(ttc_ctx *c, bool append_row) {
    k_reduce<<<1, 256, 0, c->stream>>>(reduce_pack(c, append_row));
}

// ---------------------------------------------------------------------------
// RESPA (integrate_mts.c:38-140) and Nose-Hoover chain (integrate_nhc.c:40-161)

namespace {

__global__ void __launch_bounds__(256)
k_combine_to(int n, int mask, const double4 *fpair, const double4 *g4, double4 *out, const double *partials,
             int maxb, int nbg, long long *fixj) {
    __shared__ double mean[3];
    if (mask & 4) ttc_mean_pme(partials, maxb, nbg, n, mean);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[s] = ttc_combine(s, mask, fpair, g4, mean, -1, fixj, n);
}

// v = A v + B (100/m) F; the first half kick of integrate_vverlet.c:57-59 is
// (A, B) = (vscal, vscal dt/2), the second (1, vscal dt/2)
__global__ void __launch_bounds__(256)
k_vkick(int n, double4 *vel4, const double4 *F, double A, double B, int check, const double4 *pos4,
        const double4 *anchor, const double *pcheck, int ke, double *partials, int maxb, int clear_flag,
        int *status) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear_flag && s == 0) status[ST_NLFLAG] = 0;
    bool ok = s < n;
    double4 v = make_double4(0, 0, 0, 1);
    if (ok) {
        v = vel4[s];
        double4 f = F[s];
        v.x = A * v.x + B * v.w * f.x;
        v.y = A * v.y + B * v.w * f.y;
        v.z = A * v.z + B * v.w * f.z;
        vel4[s] = v;
        if (check) {
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
    }
    if (ke) ken_block(blockIdx.x, 100.0 / v.w, v, partials, maxb, ok);
}

__device__ __forceinline__ double pbc1m(double x, double l) {
    double h = 0.5 * l;
    for (int it = 0; it < 64 && (x < -h || x > h); it++) x += (x < -h) ? l : -l;
    return x;
}

// integrate_mts.c:75-105: short half kick, drift, pbc
__global__ void __launch_bounds__(256)
k_vkick_drift(int n, double4 *pos4, double4 *vel4, const double4 *F, double A, double B, double dt, double lx,
              double ly, double lz) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double4 v = vel4[s], f = F[s], p = pos4[s];
    v.x = A * v.x + B * v.w * f.x;
    v.y = A * v.y + B * v.w * f.y;
    v.z = A * v.z + B * v.w * f.z;
    p.x = pbc1m(p.x + dt * v.x, lx);
    p.y = pbc1m(p.y + dt * v.y, ly);
    p.z = pbc1m(p.z + dt * v.z, lz);
    vel4[s] = v;
    pos4[s] = p;
}

struct NhcArgs {
    double nfree, dt;
    double w[8];
    int M, Nc, Nys, nbk;
};

// integrate_nhc.c:40-161 on one lane: kinetic energy from the kick/kinet
// per-block partials (fixed order), chain update, velocity scale ->
// st[TTC_NHC_SCALE]; target T and mass Q are read from st[TTC_NHC_T/Q]
__device__ __forceinline__ void nhc_body(const NhcArgs &a, const double *partials, int maxb, double *st) {
    __shared__ double kin;
    if (threadIdx.x < 64) {
        double t = 0.0;
        for (int b = threadIdx.x; b < a.nbk; b += 64)
            t += partials[(size_t)EV_KE0 * maxb + b] + partials[(size_t)EV_KE2 * maxb + b] +
                 partials[(size_t)EV_KE5 * maxb + b];
        t = ttc_wave_sum(t);
        if (threadIdx.x == 0) kin = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double *zr = st, *zv = st + a.M, *zg = st + 2 * a.M;
    int M = a.M, M1 = M - 1, M2 = M - 2;
    double T = st[TTC_NHC_T], Q = st[TTC_NHC_Q];
    double NkT = a.nfree * TTC_KB * T * 1E-4;
    double kT = TTC_KB * T * 1E-4;
    double akin = 2.0 * kin * 1000.0 * 1E-4 / TTC_NA;
    double scale = 1.0;
    for (int h = 0; h < a.Nc; h++)
        for (int i = 0; i < a.Nys; i++) {
            double dt = a.dt * a.w[i] / (double)a.Nc;
            double dt2 = dt / 2.0, dt4 = dt / 4.0, dt8 = dt / 8.0, AA;
            zg[0] = (akin - NkT) / Q;
            zv[M1] += zg[M1] * dt4;
            for (int j = 0; j < M1; j++) {
                AA = exp(-dt8 * zv[M1 - j]);
                zv[M2 - j] = zv[M2 - j] * AA * AA + dt4 * zg[M2 - j] * AA;
            }
            AA = exp(-dt2 * zv[0]);
            scale = scale * AA;
            for (int j = 0; j < M; j++) zr[j] += zv[j] * dt2;
            zg[0] = (scale * scale * akin - NkT) / Q;
            for (int j = 0; j < M - 1; j++) {
                AA = exp(-dt8 * zv[j + 1]);
                zv[j] = zv[j] * AA * AA + dt4 * zg[j] * AA;
                zg[j + 1] = (Q * zv[j] * zv[j] - kT) / Q;
            }
            zv[M - 1] += zg[M - 1] * dt4;
        }
    st[TTC_NHC_SCALE] = scale;
}

struct NhcPack {
    NhcArgs a;
    const double *partials;
    int maxb;
    double *st;
};

__global__ void k_nhc(NhcPack p) { nhc_body(p.a, p.partials, p.maxb, p.st); }
__global__ void kb_nhc(const NhcPack *pk, const TtcBmap *map) {
    const NhcPack &p = pk[map[blockIdx.x].rep];
    nhc_body(p.a, p.partials, p.maxb, p.st);
}

__device__ __forceinline__ void vscale_body(int bid, int n, double4 *vel4, const double *scale) {
    int s = bid * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double sc = *scale;
    double4 v = vel4[s];
    v.x *= sc; v.y *= sc; v.z *= sc;
    vel4[s] = v;
}

struct VscalePack {
    int n;
    double4 *vel4;
    const double *scale;
};

__global__ void __launch_bounds__(256) k_vscale(VscalePack a) { vscale_body(blockIdx.x, a.n, a.vel4, a.scale); }
__global__ void __launch_bounds__(256) kb_vscale(const VscalePack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const VscalePack a = pk[m.rep];
    vscale_body(m.blk, a.n, a.vel4, a.scale);
}

}  // namespace

void ttc_launch_combine_to(ttc_ctx *c, int mask, double4 *out) {
    k_combine_to<<<(c->n + 255) / 256, 256, 0, c->stream>>>(c->n, mask, c->fpair, c->g4, out, c->partials,
                                                            c->nblk_max, ttc_gather_blocks(c), c->fixj);
}

void ttc_launch_vkick(ttc_ctx *c, const double4 *F, double A, double B, bool check, bool ke, bool clear_flag) {
    k_vkick<<<ttc_kick_blocks(c), 256, 0, c->stream>>>(c->n, c->vel4, F, A, B, check ? 1 : 0, c->pos4, c->anchor,
                                                       check_consts(c), ke ? 1 : 0, c->partials, c->nblk_max,
                                                       clear_flag ? 1 : 0, c->status);
}

void ttc_launch_vkick_drift(ttc_ctx *c, const double4 *F, double A, double B, double dt) {
    k_vkick_drift<<<(c->n + 255) / 256, 256, 0, c->stream>>>(c->n, c->pos4, c->vel4, F, A, B, dt, c->p.lx, c->p.ly,
                                                             c->p.lz);
}

static NhcPack nhc_pack(ttc_ctx *c) {
    NhcPack p;
    p.a.nfree = c->nhc_nfree; p.a.dt = c->p.dt;
    for (int i = 0; i < 8; i++) p.a.w[i] = c->nhc_w[i];
    p.a.M = c->nhc_M; p.a.Nc = c->nhc_Nc; p.a.Nys = c->nhc_Nys; p.a.nbk = ttc_kick_blocks(c);
    p.partials = c->partials;
    p.maxb = c->nblk_max;
    p.st = c->nhc;
    return p;
}

void ttc_launch_nhc(ttc_ctx *c) {
    k_nhc<<<1, 64, 0, c->stream>>>(nhc_pack(c));
}

void ttc_launch_vscale(ttc_ctx *c) {
    k_vscale<<<(c->n + 255) / 256, 256, 0, c->stream>>>(VscalePack{c->n, c->vel4, c->nhc + TTC_NHC_SCALE});
}

void ttc_launch_vmul(ttc_ctx *c) {
    k_vscale<<<(c->n + 255) / 256, 256, 0, c->stream>>>(VscalePack{c->n, c->vel4, c->nhc + TTC_NHC_VMUL});
}

// ---- batched stages of the NHC step (ttc_batch.hip) ------------------------

int ttc_bstage_kinet(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<KinetPack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) { pk.push_back(kinet_pack(c)); nb.push_back(ttc_kick_blocks(c)); }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_kinet(const TtcStage &st, hipStream_t s) {
    kb_kinet<<<st.nblk, 256, 0, s>>>((const KinetPack *)st.pk, st.map);
}

int ttc_bstage_kick(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<KickPack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) { pk.push_back(kick_pack(c, 1.0)); nb.push_back(ttc_kick_blocks(c)); }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_kick(const TtcStage &st, bool energy, hipStream_t s) {
    if (energy) kb_kick<true><<<st.nblk, 256, 0, s>>>((const KickPack *)st.pk, st.map);
    else kb_kick<false><<<st.nblk, 256, 0, s>>>((const KickPack *)st.pk, st.map);
}

int ttc_bstage_nhc(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<NhcPack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) { pk.push_back(nhc_pack(c)); nb.push_back(1); }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_nhc(const TtcStage &st, hipStream_t s) {
    kb_nhc<<<st.nblk, 64, 0, s>>>((const NhcPack *)st.pk, st.map);
}

int ttc_bstage_vscale(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<VscalePack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) {
        pk.push_back(VscalePack{c->n, c->vel4, c->nhc + TTC_NHC_SCALE});
        nb.push_back((c->n + 255) / 256);
    }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_vscale(const TtcStage &st, hipStream_t s) {
    kb_vscale<<<st.nblk, 256, 0, s>>>((const VscalePack *)st.pk, st.map);
}

int ttc_bstage_reduce(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<ReducePack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) { pk.push_back(reduce_pack(c, true)); nb.push_back(1); }
    return ttc_stage_make(st, pk, nb, 0);
}
void ttc_blaunch_reduce(const TtcStage &st, hipStream_t s) {
    kb_reduce<<<st.nblk, 256, 0, s>>>((const ReducePack *)st.pk, st.map);
}
