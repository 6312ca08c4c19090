// ttc_pair.hip -- LJ + SPME real-space pair kernel (replaces the pair loop
// of cnonbond_nblist.c:134-523, STYPE + S14 + SPME branch).
//
// One lane per sorted site, one workgroup per ELL group of 64 sites.  Each
// step of the row loop reads one entry per lane from the group's current
// 16-entry list tile (4 KB, shared by the group's waves).  The
// cutoff test and minimum image run in FP64 on FP64 positions (the
// half-list kernel: on 32-bit fixed-point coordinates, with the FP64
// positions deciding inside a band around rc^2), so exactly the reference's
// pairs are evaluated; the LJ, erfc and exp arithmetic runs in FP32; each
// site's force and the energies/virial accumulate in FP64.
// Two list formats:
//   * full lists (both directions; API calls, RESPA loops, small contexts):
//     no scatter to j, every site's force is a fixed-order sum;
//   * half lists (velocity-Verlet and NHC loops, replica batches;
//     k_pair_half, kb_pair_half): each pair once, the partner's force added
//     as int64 fixed point into LDS window slots and flushed with integer
//     atomics, so the result does not depend on order.
#include "ttc_common.h"

namespace {

// the half-list kernel keeps four waves per SIMD (at most 128 VGPRs): at
// three it runs 40 % slower on chains-100k
#ifndef HALF_WPE
#define HALF_WPE __attribute__((amdgpu_waves_per_eu(4)))
#endif
#ifndef HALF_AHEAD
#define HALF_AHEAD 2   // half-list kernel: entries whose gathers are in flight ahead of the arithmetic
#endif
#ifndef PAIR_AHEAD
#define PAIR_AHEAD 2   // entries whose site gathers are in flight ahead of the arithmetic
#endif

struct PairConst {
    double lx, ly, lz, hx, hy, hz, rc2;
    double ilx, ily, ilz;     // 1/L per axis
    double erfc_rc_over_rc;   // erfc(kappa rc)/rc with the reference's erfc, FP64
    float rc2f, kappa, kpi2;  // kpi2 = 2 kappa / sqrt(pi)
    float half_kappa;         // kappa / 2
    float mk2l2;              // -kappa^2 log2(e)
    float qfac;               // QFACTOR
    float inv_epsrf;
    float e14, one_m_e14;
    int ntypes;
    // half-list kernel: displacements from 32-bit fixed-point coordinates
    // (units L / 2^32 per axis), and the band of r^2 around rc^2 inside
    // which their r^2 cannot decide the cutoff (the FP64 positions do)
    double ux, uy, uz;
    double rc2_lo, rc2_hi;
};

// Numerical Recipes erfcc (ewald.c:817-835) in FP32 at z = kappa r >= 0:
// erfc(z) = t exp(p(t) - z^2), t = 1 / (1 + z/2), with p the reference's
// degree-9 polynomial.  Evaluated with base-2 exponentials: the polynomial's
// coefficients carry the factor log2(e), and -z^2 log2(e) = mk2l2 r^2.
// Returns erfc(z); *ez2 receives exp(-z^2).
#define L2E 1.44269504088896341f
__device__ __forceinline__ float nr_erfc(float r, float r2, const PairConst &pc, float *ez2) {
    float t = __builtin_amdgcn_rcpf(fmaf(pc.half_kappa, r, 1.0f));
    float p = (L2E * -1.26551223f) + t * ((L2E * 1.00002368f) + t * ((L2E * 0.37409196f) +
              t * ((L2E * 0.09678418f) + t * ((L2E * -0.18628806f) + t * ((L2E * 0.27886807f) +
              t * ((L2E * -1.13520398f) + t * ((L2E * 1.48851587f) + t * ((L2E * -0.82215223f) +
              t * (L2E * 0.17087277f)))))))));
    float a = pc.mk2l2 * r2;
    *ez2 = __builtin_amdgcn_exp2f(a);
    return t * __builtin_amdgcn_exp2f(p + a);
}
#undef L2E


struct PairAcc {
    double fx, fy, fz;
    float tx, ty, tz;    // this tile's force, added to fx, fy, fz once per 16 entries
    double elj, eljs, er, ers, ei, w0, w1, w2, w3, w4, w5;
    double hx, hy, hz;   // CONFIGT: this site's d2U/dx2, dy2, dz2 (pair terms)
    float fmax;          // half lists: largest |force / r| (range check of the accumulators)
};

// One list entry of site i (cnonbond_nblist.c:150-519 for one pair).
struct RangeSel {
    const int *orig;
    int lb, ub;
};

// RANGE: keep only pairs with a site in [lb,ub) of the original numbering
// (enbnd_mc.c:98-119); i_in says whether the row's own site is in range.
// site gathered for entry e (site 0 for padding, which is skipped)
__device__ __forceinline__ int entry_site(uint32_t e) { return e == TTC_PAD ? 0 : (int)(e & TTC_JMASK); }

// pj: position of the entry's site, gathered by the caller (ahead of time,
// so the gathers of later entries overlap this one); its type is in the entry
template <bool ENERGY, bool RANGE = false, bool HES = false>
__device__ __forceinline__ void pair_entry(uint32_t e, const double4 &pj, const double4 &pi, float qi,
                                           const float2 *ljrow, const float2 *ljrow14, const PairConst &pc, PairAcc &a,
                                           const RangeSel &rs = RangeSel{}, bool i_in = true) {
    if (e == TTC_PAD) return;
    int j = (int)(e & TTC_JMASK);
    int tj = (int)((e >> TTC_TSHIFT) & 63u);
    uint32_t cls = e >> 30;
    // minimum image in FP64 as d - L rint(d/L): for a list pair |d| is far
    // from L/2, so this wraps exactly when the reference's d > L/2 test does,
    // and the FMA gives the reference's d -= L rounded once
    double ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
    ex = fma(-pc.lx, rint(ex * pc.ilx), ex);
    ey = fma(-pc.ly, rint(ey * pc.ily), ey);
    ez = fma(-pc.lz, rint(ez * pc.ilz), ez);
    // the reference's cutoff test on r^2 summed op by op (5 FP64 operations,
    // no band or second form to select between)
    const double dr2 = ttc_r2_exact(ex, ey, ez);
    if (!(dr2 <= pc.rc2)) return;
    const float dx = (float)ex, dy = (float)ey, dz = (float)ez;
    const float r2 = (float)dr2;
    if (RANGE && !i_in) {
        int oj = rs.orig[j];
        if (oj < rs.lb || oj >= rs.ub) return;
    }
    float rinv = __builtin_amdgcn_rsqf(r2);
    float rinv2 = rinv * rinv;
    float r = r2 * rinv;
    const bool one4 = cls == 2;
    float2 lj = ljrow[tj];
    if (__builtin_expect(one4, 0)) lj = ljrow14[tj];   // 1-4 parameters only for 1-4 pairs
    const float eps = cls == 1 ? 0.0f : lj.x;
    const float sg2 = lj.y;
    float sr2 = sg2 * rinv2;
    float sr6 = sr2 * sr2 * sr2;
    float sr12 = sr6 * sr6;
    float force = 12.0f * eps * rinv2 * (sr12 - sr6);
    float qqab = qi * (float)pj.w;   // qi carries QFACTOR
    float qq = (cls == 0 ? qqab : (cls == 2 ? pc.e14 * qqab : 0.0f)) * pc.inv_epsrf;
    float ez2;
    float erfckr = nr_erfc(r, r2, pc, &ez2);
    float kpid = pc.kpi2 * r * ez2;
    float rinv3 = rinv * rinv2;
    force += qq * rinv3 * (erfckr + kpid);
    float qi2 = cls == 2 ? pc.one_m_e14 * qqab : qqab;
    if (cls) force -= qi2 * rinv3 * ((1.0f - erfckr) - kpid);
    if (HES) {
        // Hessian diagonal of this site, one visit of the pair (the full
        // list visits it from both sites, giving the reference's factor 2):
        // cnonbond_nblist2.c:286-297 (LJ), 403-410 (real), 433-440 (excluded)
        float fdx = dx, fdy = dy, fdz = dz;
        float flj = 12.0f * eps * rinv2 * (sr12 - sr6);
        float term2 = 12.0f * eps * rinv2 * rinv2 * (14.0f * sr12 - 8.0f * sr6);
        float kpi = 0.5f * pc.kpi2;
        float rinv4 = rinv2 * rinv2, rinv5 = rinv4 * rinv;
        float e4 = 4.0f * kpi * pc.kappa * pc.kappa * rinv2 * ez2 + 6.0f * kpi * rinv4 * ez2;
        float h1 = erfckr * rinv3 + 2.0f * kpi * rinv2 * ez2;
        float h2 = -e4 - 3.0f * rinv5 * erfckr;
        float ax = term2 * fdx * fdx - flj - qq * (h1 + fdx * fdx * h2);
        float ay = term2 * fdy * fdy - flj - qq * (h1 + fdy * fdy * h2);
        float az = term2 * fdz * fdz - flj - qq * (h1 + fdz * fdz * h2);
        if (cls) {
            float om = 1.0f - erfckr;
            float g1 = om * rinv3 - 2.0f * kpi * rinv2 * ez2;
            float g2 = -e4 + 3.0f * rinv5 * om;
            ax += qi2 * (g1 - fdx * fdx * g2);
            ay += qi2 * (g1 - fdy * fdy * g2);
            az += qi2 * (g1 - fdz * fdz * g2);
        }
        a.hx += (double)ax;
        a.hy += (double)ay;
        a.hz += (double)az;
    }
    float fxf = force * dx, fyf = force * dy, fzf = force * dz;
    a.tx += fxf;
    a.ty += fyf;
    a.tz += fzf;
    if (ENERGY) {
        float sr2s = sg2 / pc.rc2f;
        float sr6s = sr2s * sr2s * sr2s;
        a.elj += (double)(eps * (sr12 - 2.0f * sr6));
        a.eljs += (double)(eps * (sr6s * sr6s - 2.0f * sr6s));
        a.er += (double)(qq * erfckr * rinv);
        a.ers += (double)qq * pc.erfc_rc_over_rc;
        if (cls) a.ei += (double)(qi2 * (1.0f - erfckr) * rinv);
        a.w0 += (double)(fxf * dx);
        a.w1 += (double)(fyf * dx);
        a.w2 += (double)(fyf * dy);
        a.w3 += (double)(fzf * dx);
        a.w4 += (double)(fzf * dy);
        a.w5 += (double)(fzf * dz);
    }
}

// One entry of a half list, without branches around the pair terms: lanes
// whose entry is padding or beyond rc compute on a harmless r^2 and their
// force is selected to zero, so the wave runs one straight sequence per
// entry (the exec-mask branches and the zeroed values around them cost more
// than the arithmetic they skip).  The partner's force goes to its window
// slot as int64 2^-24 units: -f 2^24 is exact in FP32, its FP64 image plus
// 1.5 2^52 (from an SGPR) rounds it to an integer held in the low mantissa
// bits, and removing the magic's high word leaves the two's-complement value.
// Per entry class (e >> 30: 0 plain, 1 excluded, 2 one-four, 3 padding) one
// LDS record replaces the class selects: {eps factor, charge factor of the
// real-space term (1/eps_RF, E14FAC/eps_RF for 1-4 pairs), charge factor of
// the excluded-pair correction (1, 1 - E14FAC), offset of the 1-4 LJ table
// in float2 entries (int bits)}; padding has every factor zero.
template <bool ENERGY>
__device__ __forceinline__ void pair_entry_half(uint32_t e, const int4 &pj, int j, const int4 &pi, int s, float qi,
                                                const double4 *__restrict__ pos4, const float2 *ljrow,
                                                const float4 *ctab, const PairConst &pc, PairAcc &a,
                                                unsigned long long *jacc, int wcap, double magic) {
    const int tj = (int)((e >> TTC_TSHIFT) & 63u);
    const uint32_t cls = e >> 30;
    // i - j from the fixed-point coordinates: the 32-bit difference wraps
    // modulo the box, which is the minimum image (cnonbond_nblist.c:166-173);
    // in FP64 it is exact up to the 2^-32 L quantization of the coordinates
    const double ex = (double)(int)((unsigned)pi.x - (unsigned)pj.x) * pc.ux;
    const double ey = (double)(int)((unsigned)pi.y - (unsigned)pj.y) * pc.uy;
    const double ez = (double)(int)((unsigned)pi.z - (unsigned)pj.z) * pc.uz;
    const double dr2 = fma(ex, ex, fma(ey, ey, ez * ez));
    const bool real = e != TTC_PAD;
    bool in = real & (dr2 < pc.rc2_lo);
    if (real & !(dr2 < pc.rc2_lo) & !(dr2 > pc.rc2_hi)) {
        // near rc^2 (the band: 1e-8 plus the coordinates' quantization):
        // the reference's own FP64 test on the FP64 positions decides
        // (cnonbond_nblist.c:174-176, the sum rounded op by op)
        const double4 P = pos4[s], Q = pos4[j];
        double fx = P.x - Q.x, fy = P.y - Q.y, fz = P.z - Q.z;
        fx = fma(-pc.lx, rint(fx * pc.ilx), fx);
        fy = fma(-pc.ly, rint(fy * pc.ily), fy);
        fz = fma(-pc.lz, rint(fz * pc.ilz), fz);
        in = ttc_r2_exact(fx, fy, fz) <= pc.rc2;
    }
    const float dx = (float)ex, dy = (float)ey, dz = (float)ez;
    // lanes out of range take the padding record (every factor zero, so
    // their force is zero) and compute on r^2 = 4, an inline constant; the
    // record's index depends on the cutoff test, which also keeps its load
    // next to its use
    const float4 cf = ctab[in ? cls : 3u];
    const float r2 = in ? (float)dr2 : 4.0f;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    const float r = r2 * rinv;
    const float2 lj = ljrow[tj + __float_as_int(cf.w)];
    const float eps = lj.x * cf.x;
    const float sr2 = lj.y * rinv2;
    const float sr6 = sr2 * sr2 * sr2;
    const float sr12 = sr6 * sr6;
    float force = 12.0f * eps * rinv2 * (sr12 - sr6);
    const float qqab = qi * __int_as_float(pj.w);
    const float qq = qqab * cf.y;
    float ez2;
    const float erfckr = nr_erfc(r, r2, pc, &ez2);
    const float kpid = pc.kpi2 * r * ez2;
    const float rinv3 = rinv * rinv2;
    force += qq * rinv3 * (erfckr + kpid);
    const float qi2 = qqab * cf.z;
    force -= qi2 * rinv3 * ((1.0f - erfckr) - kpid);
    const float fxf = force * dx, fyf = force * dy, fzf = force * dz;
    a.tx += fxf;
    a.ty += fyf;
    a.tz += fzf;
    // |component| <= |force| r <= |force| rc: the range check bounds |force|
    a.fmax = fmaxf(a.fmax, fabsf(force));
    if (in) {
        const int slot = (int)(e & TTC_JMASK);
        auto fix = [&](float f) {
            const double t = (double)(f * -16777216.0f) + magic;
            return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000LL);
        };
        unsigned long long *acc3 = jacc + 3 * slot;
        atomicAdd(&acc3[0], fix(fxf));
        atomicAdd(&acc3[1], fix(fyf));
        atomicAdd(&acc3[2], fix(fzf));
    }
    if (ENERGY && in) {
        const float sr2s = lj.y / pc.rc2f;
        const float sr6s = sr2s * sr2s * sr2s;
        a.elj += (double)(eps * (sr12 - 2.0f * sr6));
        a.eljs += (double)(eps * (sr6s * sr6s - 2.0f * sr6s));
        a.er += (double)(qq * erfckr * rinv);
        a.ers += (double)qq * pc.erfc_rc_over_rc;
        a.ei += (double)(qi2 * (1.0f - erfckr) * rinv);
        a.w0 += (double)(fxf * dx);
        a.w1 += (double)(fyf * dx);
        a.w2 += (double)(fyf * dy);
        a.w3 += (double)(fzf * dx);
        a.w4 += (double)(fzf * dy);
        a.w5 += (double)(fzf * dz);
    }
}

// One workgroup per ELL group of 64 sites; its WAVES waves take the row
// slices k = wave, wave + WAVES, ... (the waves together sweep one 16-entry
// tile of every row at a time) and their partial forces are summed in a
// fixed wave order.  More
// waves per group shorten each group's dependent load chain (small boxes,
// few groups); fewer suit many groups.  ttc_tune picks WAVES per context.
template <bool ENERGY, bool RANGE, bool HES, int WAVES>
__device__ __forceinline__ void pair_body(int grp, int n, int cap, const uint32_t *__restrict__ nbl,
                                          const int *__restrict__ glen, const double4 *__restrict__ pos4,
                                          const int *__restrict__ type, const float4 *__restrict__ ljtab,
                                          const PairConst &pc, double4 *__restrict__ fpair,
                                          double *__restrict__ partials, int maxb, const RangeSel &rs) {
    // LJ pair table in two halves: {eps, sig^2} and the 1-4 {eps, sig^2}
    extern __shared__ float2 slj[];
    __shared__ double sred[WAVES][64][3];
    __shared__ double ered[WAVES][14];
    int nt2 = pc.ntypes * pc.ntypes;
    for (int k = threadIdx.x; k < nt2; k += blockDim.x) {
        const float4 v = ljtab[k];
        slj[k] = make_float2(v.x, v.y);
        slj[nt2 + k] = make_float2(v.z, v.w);
    }
    __syncthreads();

    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int s = grp * 64 + lane;
    PairAcc a = {};
    int len = glen[grp];
    double4 pi = s < n ? pos4[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    int ti = s < n ? type[s] : 0;
    const float2 *ljrow = slj + ti * pc.ntypes, *ljrow14 = slj + nt2 + ti * pc.ntypes;
    float qi = (float)pi.w * pc.qfac;
    // this lane's 64-byte segment of tile 0; tile t is 1024 words further
    const uint4 *seg = (const uint4 *)(nbl + (size_t)grp * cap * 64 + (size_t)lane * 16);
    bool i_in = true;
    if (RANGE) {
        int oi = s < n ? rs.orig[s] : -1;
        i_in = oi >= rs.lb && oi < rs.ub;
    }
    // wave w takes the whole tiles w, w + WAVES, ...: each lane reads its
    // 16 entries of a tile as four 16-byte loads, the next tile's while the
    // current one is evaluated (rows are padded to a multiple of 16)
    int ntile = (len + 15) >> 4;
    int t = wave;
    uint4 q0, q1, q2, q3;
    if (t < ntile) {
        const uint4 *p = seg + (size_t)t * 256;
        q0 = p[0]; q1 = p[1]; q2 = p[2]; q3 = p[3];
    }
    for (; t < ntile; t += WAVES) {
        uint32_t e[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                          q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
        // gathers run PAIR_AHEAD entries ahead of the arithmetic
        double4 pj[16];
#pragma unroll
        for (int q = 0; q < PAIR_AHEAD; q++) pj[q] = pos4[entry_site(e[q])];
        if (t + WAVES < ntile) {
            const uint4 *p = seg + (size_t)(t + WAVES) * 256;
            q0 = p[0]; q1 = p[1]; q2 = p[2]; q3 = p[3];
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (q + PAIR_AHEAD < 16) pj[q + PAIR_AHEAD] = pos4[entry_site(e[q + PAIR_AHEAD])];
            pair_entry<ENERGY, RANGE, HES>(e[q], pj[q], pi, qi, ljrow, ljrow14, pc, a, rs, i_in);
        }
        // the tile's FP32 force into the FP64 sum
        a.fx += (double)a.tx;
        a.fy += (double)a.ty;
        a.fz += (double)a.tz;
        a.tx = a.ty = a.tz = 0.0f;
    }

    sred[wave][lane][0] = a.fx;
    sred[wave][lane][1] = a.fy;
    sred[wave][lane][2] = a.fz;
    if (ENERGY) {
        const int nq = HES ? 14 : 11;
        double v[14] = {a.elj, a.eljs, a.er, a.ers, a.ei, a.w0, a.w1, a.w2, a.w3, a.w4, a.w5, a.hx, a.hy, a.hz};
        for (int q = 0; q < nq; q++) {
            double t = ttc_wave_sum(v[q]);
            if (lane == 0) ered[wave][q] = t;
        }
    }
    __syncthreads();
    if (wave == 0 && s < n) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int w = 0; w < WAVES; w++) {
            fx += sred[w][lane][0];
            fy += sred[w][lane][1];
            fz += sred[w][lane][2];
        }
        fpair[s] = make_double4(fx, fy, fz, 0.0);
    }
    if (ENERGY && threadIdx.x < 11) {
        // each pair was visited from both of its sites: halve
        double t = 0.0;
        for (int w = 0; w < WAVES; w++) t += ered[w][threadIdx.x];
        partials[(size_t)(EV_LJ + threadIdx.x) * maxb + grp] = 0.5 * t;
    }
    if (HES && threadIdx.x >= 11 && threadIdx.x < 14) {
        // per-site diagonal terms: already one per visit, no halving
        double t = 0.0;
        for (int w = 0; w < WAVES; w++) t += ered[w][threadIdx.x];
        partials[(size_t)(EV_PHX + threadIdx.x - 11) * maxb + grp] = t;
    }
}

struct PairPack {
    int n, cap;
    const uint32_t *nbl;
    const int *glen;
    const double4 *pos4;
    const int *type;
    const float4 *ljtab;
    PairConst pc;
    double4 *fpair;
    double *partials;
    int maxb;
    RangeSel rs;
    int *stepst = nullptr;   // MD step: count the step, clear the consumed trigger flag
};

template <bool ENERGY, bool RANGE, bool HES, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_pair(PairPack a) {
    if (a.stepst && blockIdx.x == 0 && threadIdx.x == 0) {
        a.stepst[ST_STEP] += 1;
        a.stepst[ST_NLFLAG] = 0;
    }
    pair_body<ENERGY, RANGE, HES, WAVES>(blockIdx.x, a.n, a.cap, a.nbl, a.glen, a.pos4, a.type, a.ljtab, a.pc, a.fpair,
                                         a.partials, a.maxb, a.rs);
}
template <bool ENERGY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) kb_pair(const PairPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const PairPack a = pk[m.rep];
    pair_body<ENERGY, false, false, WAVES>(m.blk, a.n, a.cap, a.nbl, a.glen, a.pos4, a.type, a.ljtab, a.pc, a.fpair,
                                           a.partials, a.maxb, a.rs);
}

// Half list: one workgroup per ELL group of 64 sites as above; entries
// hold window slots.  The group's window (the distinct partners j, win[g])
// is staged in LDS together with three int64 accumulators per slot; each
// pair's force is added to i in registers and, negated, to j's slot.  After
// the row loop the slots are added to fixj[3][n] with one global 64-bit
// atomic per slot and component (2^-24 kJ/mol/A units: integer adds, so the
// total does not depend on the order).  Energies and the virial count each
// pair once.
struct PairHalfPack {
    int n, cap, wcap;
    const uint32_t *nbl;
    const int *glen;
    const int *win, *wlen;
    const double4 *pos4;
    const int4 *posq;      // fixed-point coordinates and charge bits, as pos4
    const int *type;
    const float4 *ljtab;
    PairConst pc;
    double4 *fpair;
    long long *fixj;
    double *partials;
    int maxb;
    int *status;
    int step;   // MD step: count the step, clear the consumed trigger flag, stop after an overflow
    double magic;   // 1.5 2^52 (rounding constant of the partner forces)
};

template <bool ENERGY, int WAVES>
__device__ __forceinline__ void pair_half_body(int grp, const PairHalfPack &a) {
    extern __shared__ unsigned long long hsm[];
    const int wcap = a.wcap;
    unsigned long long *jacc = hsm;                       // [wcap][3]
    float2 *slj = (float2 *)(hsm + 3 * (size_t)wcap);     // LJ halves
    const int nt2 = a.pc.ntypes * a.pc.ntypes;
    int *swin = (int *)(slj + 2 * nt2);                   // [wcap]
    __shared__ double sred[WAVES][64][3];
    __shared__ double ered[WAVES][11];
    __shared__ float sfmax[WAVES];
    __shared__ float4 ctab[4];
    if (threadIdx.x == 0) {
        ctab[0] = make_float4(1.0f, a.pc.inv_epsrf, 0.0f, __int_as_float(0));
        ctab[1] = make_float4(0.0f, 0.0f, 1.0f, __int_as_float(0));
        ctab[2] = make_float4(1.0f, a.pc.e14 * a.pc.inv_epsrf, a.pc.one_m_e14, __int_as_float(nt2));
        ctab[3] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
    }
    const int W = min(a.wlen[grp], wcap);
    const int *gwin = a.win + (size_t)grp * wcap;
    // slots past the window (only after an overflow, whose step is redone)
    // read site 0 and are not flushed
    for (int k = threadIdx.x; k < wcap; k += blockDim.x) {
        swin[k] = k < W ? gwin[k] : 0;
        jacc[3 * k] = 0ull;
        jacc[3 * k + 1] = 0ull;
        jacc[3 * k + 2] = 0ull;
    }
    for (int k = threadIdx.x; k < nt2; k += blockDim.x) {
        const float4 v = a.ljtab[k];
        slj[k] = make_float2(v.x, v.y);
        slj[nt2 + k] = make_float2(v.z, v.w);
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = a.n;
    const int s = grp * 64 + lane;
    PairAcc acc = {};
    const int len = a.glen[grp];
    const int si = s < n ? s : 0;
    const int4 pi = a.posq[si];
    const int ti = s < n ? a.type[s] : 0;
    const float2 *ljrow = slj + ti * a.pc.ntypes;   // the 1-4 row is nt2 entries further
    const float qi = __int_as_float(pi.w) * a.pc.qfac;
    const uint4 *seg = (const uint4 *)(a.nbl + (size_t)grp * a.cap * 64 + (size_t)lane * 16);
    const int ntile = (len + 15) >> 4;
    int t = wave;
    uint4 q0, q1, q2, q3;
    if (t < ntile) {
        const uint4 *p = seg + (size_t)t * 256;
        q0 = p[0]; q1 = p[1]; q2 = p[2]; q3 = p[3];
    }
    for (; t < ntile; t += WAVES) {
        const uint32_t e[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
        // partner sites of the tile from the window (padding: site 0)
        int js[16];
#pragma unroll
        for (int q = 0; q < 16; q++) js[q] = e[q] == TTC_PAD ? 0 : swin[e[q] & TTC_JMASK];
        int4 pj[16];
#pragma unroll
        for (int q = 0; q < HALF_AHEAD; q++) pj[q] = a.posq[js[q]];
        if (t + WAVES < ntile) {
            const uint4 *p = seg + (size_t)(t + WAVES) * 256;
            q0 = p[0]; q1 = p[1]; q2 = p[2]; q3 = p[3];
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (q + HALF_AHEAD < 16) pj[q + HALF_AHEAD] = a.posq[js[q + HALF_AHEAD]];
            pair_entry_half<ENERGY>(e[q], pj[q], js[q], pi, si, qi, a.pos4, ljrow, ctab, a.pc, acc, jacc, wcap,
                                    a.magic);
        }
        acc.fx += (double)acc.tx;
        acc.fy += (double)acc.ty;
        acc.fz += (double)acc.tz;
        acc.tx = acc.ty = acc.tz = 0.0f;
    }

    sred[wave][lane][0] = acc.fx;
    sred[wave][lane][1] = acc.fy;
    sred[wave][lane][2] = acc.fz;
    float fm = acc.fmax;
    for (int o = 32; o > 0; o >>= 1) fm = fmaxf(fm, __shfl_xor(fm, o, 64));
    if (lane == 0) sfmax[wave] = fm;
    if (ENERGY) {
        double v[11] = {acc.elj, acc.eljs, acc.er, acc.ers, acc.ei, acc.w0, acc.w1, acc.w2, acc.w3, acc.w4, acc.w5};
        for (int q = 0; q < 11; q++) {
            double tq = ttc_wave_sum(v[q]);
            if (lane == 0) ered[wave][q] = tq;
        }
    }
    __syncthreads();
    // partners' forces to the global accumulators
    long long *fixj = a.fixj;
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
        const int j = swin[k];
        const long long vx = (long long)jacc[3 * k], vy = (long long)jacc[3 * k + 1], vz = (long long)jacc[3 * k + 2];
        if (vx) atomicAdd((unsigned long long *)&fixj[j], (unsigned long long)vx);
        if (vy) atomicAdd((unsigned long long *)&fixj[n + j], (unsigned long long)vy);
        if (vz) atomicAdd((unsigned long long *)&fixj[2 * n + j], (unsigned long long)vz);
    }
    if (wave == 0 && s < n) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int w = 0; w < WAVES; w++) {
            fx += sred[w][lane][0];
            fy += sred[w][lane][1];
            fz += sred[w][lane][2];
        }
        a.fpair[s] = make_double4(fx, fy, fz, 0.0);
    }
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int w = 0; w < WAVES; w++) m = fmaxf(m, sfmax[w]);
        // beyond 2^26 kJ/mol/A a component would leave the 2^51 range of the
        // partner accumulators' rounding: |force / r| rc < 2^26
        if (!(m * m * a.pc.rc2f < 4503599627370496.0f)) atomicOr(&a.status[ST_OVERFLOW], 16);
    }
    if (ENERGY && threadIdx.x < 11) {
        double tq = 0.0;
        for (int w = 0; w < WAVES; w++) tq += ered[w][threadIdx.x];
        a.partials[(size_t)(EV_LJ + threadIdx.x) * a.maxb + grp] = tq;
    }
}

template <bool ENERGY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) HALF_WPE k_pair_half(PairHalfPack a) {
    if (a.step) {
        if (a.status[ST_OVERFLOW] & TTC_OVF_REDO) return;   // a rebuild of this loop overflowed
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.status[ST_STEP] += 1;
            a.status[ST_NLFLAG] = 0;
        }
    }
    pair_half_body<ENERGY, WAVES>(blockIdx.x, a);
}

// batched (ttc_batch.hip): block `blk` of context `rep`, its argument pack
// read in place
template <bool ENERGY, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) HALF_WPE kb_pair_half(const PairHalfPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    pair_half_body<ENERGY, WAVES>(m.blk, pk[m.rep]);
}

}  // namespace

extern double ttc_nr_erfc_host(double x);

int ttc_pair_blocks(ttc_ctx *c) { return c->G; }

static PairConst pair_const(const ttc_ctx *c) {
    PairConst pc;
    pc.lx = c->p.lx; pc.ly = c->p.ly; pc.lz = c->p.lz;
    pc.hx = 0.5 * c->p.lx; pc.hy = 0.5 * c->p.ly; pc.hz = 0.5 * c->p.lz;
    pc.rc2 = c->p.rc * c->p.rc;
    pc.ilx = 1.0 / c->p.lx; pc.ily = 1.0 / c->p.ly; pc.ilz = 1.0 / c->p.lz;
    pc.erfc_rc_over_rc = ttc_nr_erfc_host(c->p.kappa * c->p.rc) / c->p.rc;
    pc.rc2f = (float)pc.rc2;
    pc.kappa = (float)c->p.kappa;
    pc.kpi2 = (float)(2.0 * c->p.kappa / sqrt(TTC_PI));
    pc.half_kappa = (float)(0.5 * c->p.kappa);
    pc.mk2l2 = (float)(-c->p.kappa * c->p.kappa / log(2.0));
    pc.qfac = (float)TTC_QFACTOR;
    pc.inv_epsrf = (float)(1.0 / c->p.eps_rf);
    pc.e14 = (float)c->p.e14fac;
    pc.one_m_e14 = (float)(1.0 - c->p.e14fac);
    pc.ntypes = c->ntypes;
    pc.ux = c->p.lx / 4294967296.0;
    pc.uy = c->p.ly / 4294967296.0;
    pc.uz = c->p.lz / 4294967296.0;
    // the fixed-point r^2 differs from the FP64 positions' r^2 by the
    // coordinates' quantization: each coordinate is rounded to L / 2^32, so
    // a difference is within L / 2^32 of the FP64 one and |r_q^2 - r^2| <=
    // 2 r sqrt(3) Lmax 2^-32 (plus its negligible square), 2 sqrt(3) Lmax /
    // (rc 2^32) of rc^2 at r ~ rc (1.6e-8 for a 200 A box at rc = 10 A).
    // Outside twice that, plus 1e-8 for the FP64 roundings, it decides the
    // cutoff alone
    const double lmax = std::max(c->p.lx, std::max(c->p.ly, c->p.lz));
    const double band = 1e-8 + 4.0 * sqrt(3.0) * lmax / (c->p.rc * 4294967296.0);
    pc.rc2_lo = pc.rc2 * (1.0 - band);
    pc.rc2_hi = pc.rc2 * (1.0 + band);
    return pc;
}

static PairPack pair_pack(ttc_ctx *c, RangeSel rs) {
    return PairPack{c->n, c->cap, c->nbl, c->glen, c->pos4, c->type, c->ljtab, pair_const(c), c->fpair, c->partials,
                    c->nblk_max, rs};
}

template <bool ENERGY, bool RANGE, bool HES = false>
static void launch_pair_w(ttc_ctx *c, RangeSel rs, int *stepst = nullptr) {
    int blocks = ttc_pair_blocks(c);
    size_t shm = sizeof(float4) * c->ntypes * c->ntypes;
    PairPack a = pair_pack(c, rs);
    a.stepst = stepst;
#define PAIR_LAUNCH(W) k_pair<ENERGY, RANGE, HES, W><<<blocks, 64 * W, shm, c->stream>>>(a)
    switch (c->pair_waves) {
    case 4: PAIR_LAUNCH(4); break;
    case 16: PAIR_LAUNCH(16); break;
    default: PAIR_LAUNCH(8); break;
    }
#undef PAIR_LAUNCH
}

void ttc_launch_pair_range(ttc_ctx *c, int lb, int ub) {
    launch_pair_w<true, true>(c, RangeSel{c->orig, lb, ub});
}

void ttc_launch_pair_hes(ttc_ctx *c) {
    launch_pair_w<true, false, true>(c, RangeSel{});
}

size_t ttc_pair_half_shm(const ttc_ctx *c) {
    return (sizeof(unsigned long long) * 3 + sizeof(int)) * (size_t)c->wcap +
           sizeof(float2) * 2 * (size_t)c->ntypes * c->ntypes;
}

namespace {
// posq[s] = {x, y, z} as 32-bit fixed point (x 2^32 / L, wrapped modulo
// 2^32, i.e. modulo the box) and the charge's bits, from pos4
__global__ void __launch_bounds__(256) k_posq(int n, const double4 *__restrict__ pos4, int4 *__restrict__ posq,
                                              double kx, double ky, double kz) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    posq[s] = ttc_posq(pos4[s], kx, ky, kz);
}
}  // namespace

void ttc_launch_posq(ttc_ctx *c) {
    const double two32 = 4294967296.0;
    k_posq<<<(c->n + 255) / 256, 256, 0, c->stream>>>(c->n, c->pos4, c->posq, two32 / c->p.lx, two32 / c->p.ly,
                                                       two32 / c->p.lz);
}

// an MD step's drift (and its loop's rebuild) keep posq current; other
// callers refresh it from pos4 first (refresh)
static void launch_pair_half(ttc_ctx *c, bool energy, bool step, bool refresh = true) {
    if (refresh && !step) ttc_launch_posq(c);
    PairHalfPack a{c->n, c->cap, c->wcap, c->nbl, c->glen, c->win, c->wlen, c->pos4, c->posq, c->type, c->ljtab,
                   pair_const(c), c->fpair, c->fixj, c->partials, c->nblk_max, c->status, step ? 1 : 0,
                   6755399441055744.0};
    size_t shm = ttc_pair_half_shm(c);
#define PH_LAUNCH(E, W) k_pair_half<E, W><<<c->G, 64 * W, shm, c->stream>>>(a)
    if (c->pair_waves_h == 4) {
        if (energy) PH_LAUNCH(true, 4); else PH_LAUNCH(false, 4);
    } else {
        if (energy) PH_LAUNCH(true, 8); else PH_LAUNCH(false, 8);
    }
#undef PH_LAUNCH
}

// LDS above 64 KiB for the half-list kernel (gfx950: 160 KiB per CU)
int ttc_pair_half_attrs(ttc_ctx *c) {
    int shm = (int)ttc_pair_half_shm(c);
    const void *fns[] = {(const void *)k_pair_half<true, 4>, (const void *)k_pair_half<false, 4>,
                         (const void *)k_pair_half<true, 8>, (const void *)k_pair_half<false, 8>};
    for (const void *f : fns) TTC_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    return TTC_OK;
}

// the pair kernel alone on the current posq (timing loops: ttc_launch_posq
// once before them)
void ttc_launch_pair_only(ttc_ctx *c) {
    if (c->list_half) launch_pair_half(c, false, false, false);
    else launch_pair_w<false, false>(c, RangeSel{});
}

void ttc_launch_pair(ttc_ctx *c, bool energy, int /*blk_off*/) {
    if (c->list_half) launch_pair_half(c, energy, false);
    else if (energy) launch_pair_w<true, false>(c, RangeSel{});
    else launch_pair_w<false, false>(c, RangeSel{});
}

// after a drift (which wrote the fixed-point coordinates) outside the
// velocity-Verlet loop: the current list's kernel, no step counting
void ttc_launch_pair_drifted(ttc_ctx *c, bool energy) {
    if (c->list_half) launch_pair_half(c, energy, false, false);
    else if (energy) launch_pair_w<true, false>(c, RangeSel{});
    else launch_pair_w<false, false>(c, RangeSel{});
}

void ttc_launch_pair_step(ttc_ctx *c, bool energy) {
    if (c->list_half) launch_pair_half(c, energy, true);
    else if (energy) launch_pair_w<true, false>(c, RangeSel{}, c->status);
    else launch_pair_w<false, false>(c, RangeSel{}, c->status);
}

// ---- batched stage of the NHC step (ttc_batch.hip): every context's pair
// kernel with 8 waves per group (the untuned default of a context), the
// half-list kernel when every context holds a half list
int ttc_bstage_pair(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    bool half = !r.empty();
    for (ttc_ctx *c : r) half = half && c->list_half;
    std::vector<int> nb;
    size_t shm = 0;
    if (half) {
        std::vector<PairHalfPack> pk;
        for (ttc_ctx *c : r) {
            pk.push_back(PairHalfPack{c->n, c->cap, c->wcap, c->nbl, c->glen, c->win, c->wlen, c->pos4, c->posq,
                                      c->type, c->ljtab, pair_const(c), c->fpair, c->fixj, c->partials, c->nblk_max,
                                      c->status, 0, 6755399441055744.0});
            nb.push_back(c->G);
            shm = std::max(shm, ttc_pair_half_shm(c));
        }
        const void *fns[] = {(const void *)kb_pair_half<true, 8>, (const void *)kb_pair_half<false, 8>};
        for (const void *f : fns) TTC_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        int e = ttc_stage_make(st, pk, nb, shm);
        st->kind = 1;
        return e;
    }
    std::vector<PairPack> pk;
    for (ttc_ctx *c : r) {
        pk.push_back(pair_pack(c, RangeSel{}));
        nb.push_back(ttc_pair_blocks(c));
        shm = std::max(shm, sizeof(float4) * c->ntypes * c->ntypes);
    }
    int e = ttc_stage_make(st, pk, nb, shm);
    st->kind = 0;
    return e;
}

void ttc_blaunch_pair(const TtcStage &st, bool energy, hipStream_t s) {
    if (st.kind == 1) {
        if (energy) kb_pair_half<true, 8><<<st.nblk, 64 * 8, st.shm, s>>>((const PairHalfPack *)st.pk, st.map);
        else kb_pair_half<false, 8><<<st.nblk, 64 * 8, st.shm, s>>>((const PairHalfPack *)st.pk, st.map);
        return;
    }
    if (energy) kb_pair<true, 8><<<st.nblk, 64 * 8, st.shm, s>>>((const PairPack *)st.pk, st.map);
    else kb_pair<false, 8><<<st.nblk, 64 * 8, st.shm, s>>>((const PairPack *)st.pk, st.map);
}
