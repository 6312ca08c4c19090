// ttc_mc.hip -- batched Go-model Monte Carlo with Wang-Landau acceptance and
// a learned delayed-acceptance first stage (include/ttc_mc.h).
//
// One workgroup per walker runs every trial of its walker: the chain lives
// in LDS, proposals are drawn by one lane, the exact FP64 energy change is a
// block-wide sum over the bonds and pairs touching the moved beads (the
// range energy of forces_mc.c:90-300 / enbnd_mc.c:26-504), reduced in a fixed
// order, and the Wang-Landau bookkeeping (pivot.c:236-253) stays in LDS.
// The surrogate network (in -> 64 tanh -> 64 tanh -> 1) runs in the same
// workgroup with its weights in LDS; its training is a single-workgroup
// Adam loop over the logged moves.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ttc.h"
#include "../../include/ttc_mc.h"

void ttc_set_error(const std::string &msg);

#define MC_THREADS 1024
#define MC_WAVES (MC_THREADS / 64)
#define NF TTC_MC_NFEAT
#define HID TTC_MC_HID
#define NPAR (HID * NF + HID + HID * HID + HID + HID + 1)
#define MC_PI 3.141592653589793238462643383

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ttc_set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return TTC_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

struct ttc_mc {
    ttc_mc_params p;
    double *pos = nullptr;      // [nwalk][maxn][3]
    int *nbead = nullptr;       // [nwalk]
    // native contacts per walker, CSR over beads, both directions, partners
    // ascending: noff[w][maxn+1], npart[w][ncap], nr0[w][ncap]
    int *noff = nullptr, *npart = nullptr;
    double *nr0 = nullptr;
    int ncap = 0;               // CSR capacity per walker (maxn * TTC_MC_MAXNAT)
    int max_nnz = 0;            // largest CSR of the walkers set so far
    double *emin = nullptr, *ecur = nullptr;
    double *lng = nullptr;      // [nwalk][nbins]
    long long *hist = nullptr;  // [nwalk][nbins]
    uint64_t *rng = nullptr;    // [nwalk][4]
    long long *stats = nullptr; // [nwalk][TTC_MC_NSTAT]
    float *mlp = nullptr;       // [NPAR]
    float *adam = nullptr;      // [2*NPAR]
    int adam_t = 0;
    float *logx = nullptr;      // [logcap][NF]
    double *logde = nullptr;    // [logcap]
    int *logn = nullptr;
    int logcap = 0;
    hipStream_t stream = nullptr;
    std::vector<int> hn;        // host copy of bead counts
    std::vector<uint64_t> chain_sig; // per walker: hash of bead count and native contacts
    // device copy of the sampling state taken by ttc_mc_checkpoint
    double *ck_pos = nullptr, *ck_ecur = nullptr, *ck_lng = nullptr;
    long long *ck_hist = nullptr;
    uint64_t *ck_rng = nullptr;
    double t_last = 0.0;
    int dev = 0;                // device the context was created on
};

// ---------------------------------------------------------------- device

// xoshiro256** (Blackman and Vigna), one state per walker
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
__device__ __forceinline__ uint64_t xo_next(uint64_t *s) {
    uint64_t r = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return r;
}
__device__ __forceinline__ double xo_uniform(uint64_t *s) { return (double)(xo_next(s) >> 11) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ float bf16_round(float x) {
    uint32_t u = __float_as_uint(x);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return __uint_as_float(u & 0xFFFF0000u);
}

// pair term of the move features, min((sig/r)^6, 16), in FP32 (the result
// is rounded to bfloat16)
__device__ __forceinline__ float feat_term(float r2, float sig2) {
    float s = sig2 * __builtin_amdgcn_rcpf(r2);
    float v = s * s * s;
    return v < 16.0f ? v : 16.0f;
}

struct McArgs {
    int nwalk, maxn, nbins;
    double ebin, kbond, r0bond, eps, sig2, pivot_frac, theta_max, disp_max;
};

// Lane exchanges inside rows of 16 lanes by DPP (no LDS round trip): partner
// lane i^1, i^2 (quad permutations), 7-i within each 8 (half mirror) and
// 15-i within each 16 (mirror).  All 64 lanes must be active.
#define DPP_XOR1 0xB1
#define DPP_XOR2 0x4E
#define DPP_HALF_MIRROR 0x141
#define DPP_MIRROR 0x140

template <int CTRL> __device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane_f32(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Wave-wide sum (and minimum): the partners' values are combined in the
// same order in every lane (a + b = b + a), rows of 16 first, then the four
// row totals as ((r0 + r1) + (r2 + r3)); the result is wave-uniform.
__device__ __forceinline__ double wave_sum_f64(double v) {
    v += dpp_f64<DPP_XOR1>(v);
    v += dpp_f64<DPP_XOR2>(v);
    v += dpp_f64<DPP_HALF_MIRROR>(v);
    v += dpp_f64<DPP_MIRROR>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ __forceinline__ float wave_sum_f32(float v) {
    v += dpp_f32<DPP_XOR1>(v);
    v += dpp_f32<DPP_XOR2>(v);
    v += dpp_f32<DPP_HALF_MIRROR>(v);
    v += dpp_f32<DPP_MIRROR>(v);
    return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}
__device__ __forceinline__ float wave_min_f32(float v) {
    v = fminf(v, dpp_f32<DPP_XOR1>(v));
    v = fminf(v, dpp_f32<DPP_XOR2>(v));
    v = fminf(v, dpp_f32<DPP_HALF_MIRROR>(v));
    v = fminf(v, dpp_f32<DPP_MIRROR>(v));
    return fminf(fminf(readlane_f32(v, 0), readlane_f32(v, 16)), fminf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

// Block-wide FP64 sum in a fixed order.  red[] is written before the one
// barrier and read after it: callers separate consecutive calls by a barrier.
__device__ double block_sum(double v, double *red) {
    v = wave_sum_f64(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < MC_WAVES; i++) t += red[i];
    return t;
}

// 1/x to within about an ulp: hardware reciprocal refined by two Newton steps
__device__ __forceinline__ double recip64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// sin and cos of |x| <= pi/4 from their Taylor series, Horner in x^2 with
// the coefficients (-1)^k/(2k+1)! (sine, k = 8..1) and (-1)^k/(2k)! (cosine,
// k = 9..1) up to x^17 and x^18 (the first omitted terms are below 1e-19
// there); larger angles take the library routines.  The table is read from
// memory (the sampling kernel keeps it in LDS) so that the compiler does not
// pin 17 FP64 constants in registers across the trial loop.
__constant__ double kSinCosTaylor[17] = {
    -1.0 / 355687428096000.0, 1.0 / 1307674368000.0, -1.0 / 6227020800.0, 1.0 / 39916800.0,
    -1.0 / 362880.0, 1.0 / 5040.0, -1.0 / 120.0, 1.0 / 6.0,
    1.0 / 6402373705728000.0, -1.0 / 20922789888000.0, 1.0 / 87178291200.0, -1.0 / 479001600.0,
    1.0 / 3628800.0, -1.0 / 40320.0, 1.0 / 720.0, -1.0 / 24.0, 0.5};
__device__ __forceinline__ void sincos_pivot(double x, const double *tc, double *sn, double *c) {
    if (!(fabs(x) <= 0.78539816339744831)) { *sn = sin(x); *c = cos(x); return; }
    double x2 = x * x;
    double ps = tc[0];
#pragma unroll
    for (int k = 1; k < 8; k++) ps = fma(ps, x2, tc[k]);
    *sn = fma(-x * x2, ps, x);
    double pc = tc[8];
#pragma unroll
    for (int k = 9; k < 17; k++) pc = fma(pc, x2, tc[k]);
    *c = fma(-x2, pc, 1.0);
}

// Go pair energy from 1/r^2 (cgolik.c:1209-1229 form for natives,
// eps (sig/r)^12 for the rest)
__device__ __forceinline__ double go_pair_inv(double inv_r2, double r0, double eps, double sig2) {
    double s2 = (r0 > 0.0 ? r0 * r0 : sig2) * inv_r2;
    double s4 = s2 * s2, s6 = s4 * s2, s12 = s6 * s6;
    return r0 > 0.0 ? eps * (5.0 * s12 - 6.0 * (s6 * s4)) : eps * s12;
}

// E(new) - E(old) of one pair from a = r_new^2 and b = r_old^2 with one
// reciprocal: E(r) = eps (c12 (s/r^2)^6 - c10 (s/r^2)^5) with (c12, c10, s) =
// (5, 6, r0^2) for a native pair (cgolik.c:1209-1229 form) and (1, 0, sig^2)
// for the rest, so
//   E(a) - E(b) = eps [c12 s^6 (b^6 - a^6) - c10 s^5 (a b^6 - b a^6)] / (a^6 b^6).
// (a^6 b^6 stays inside the FP64 range for pair distances from 1e-3 to 1e5 A.)
__device__ __forceinline__ double go_pair_delta(double a, double b, double r0, double eps, double sig2) {
    const double a3 = a * a * a, b3 = b * b * b, a6 = a3 * a3, b6 = b3 * b3;
    const double inv = recip64(a6 * b6);
    if (r0 > 0.0) {
        const double s = r0 * r0, s2 = s * s, s5 = s2 * s2 * s, s6 = s5 * s;
        return eps * inv * (5.0 * s6 * (b6 - a6) - 6.0 * s5 * (a * b6 - b * a6));
    }
    const double s3 = sig2 * sig2 * sig2;
    return eps * (s3 * s3) * (b6 - a6) * inv;
}

// A walker's native contacts: CSR over beads (both directions), partners in
// ascending order, native distance r0 per entry.
struct Natives {
    const int *off, *part;
    const double *r0;
};

// Native distance of (b, ia) for an ascending sequence of ia: *p walks
// b's partner list, *nextp is the partner at *p (INT_MAX past the end).
__device__ __forceinline__ double native_r0(const Natives &nat, int b, int ia, int *p, int *nextp) {
    int pe = nat.off[b + 1];
    while (*nextp < ia) {
        (*p)++;
        *nextp = *p < pe ? nat.part[*p] : 0x7FFFFFFF;
    }
    return *nextp == ia ? nat.r0[*p] : 0.0;
}

// first entry of b's partner list with partner >= ia0 (binary search)
__device__ __forceinline__ void native_start(const Natives &nat, int b, int ia0, int *p, int *nextp) {
    int lo = nat.off[b], pe = nat.off[b + 1], hi = pe;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (nat.part[mid] < ia0) lo = mid + 1;
        else hi = mid;
    }
    *p = lo;
    *nextp = lo < pe ? nat.part[lo] : 0x7FFFFFFF;
}

// Energy change of walker when beads [s,e) take newp[], all other beads fixed:
// sum over bonds (i,i+1) and pairs |a-b|>=3 touching [s,e) of E(new)-E(old)
// (forces_mc(lb,ub,1) - forces_mc(lb,ub,0)).  Each thread owns a bead b and
// walks the moved beads a (the workgroup split into groups of n threads,
// group g taking a = s+g, s+g+S, ...); pairs with both beads moved are
// counted once, from the smaller index.
// rigid: [s,e) moved as a rigid body (pivot rotation), so pairs inside it
// keep their distances and their energy change is zero: only the fixed beads
// own pairs, and the workgroup is split over them alone.
__device__ double range_delta_part(const double *pos, const double *newp, int n, int s, int e, const Natives &nat,
                                   const McArgs &a, bool rigid) {
    double acc = 0.0;
    if (rigid) {
        const int m = e - s, nf = n - m;
        const int S = max(1, (int)blockDim.x / max(nf, 1));
        for (int idx = threadIdx.x; idx < nf * S; idx += blockDim.x) {
            int g = idx / nf, r = idx - g * nf;
            int b = r < s ? r : r + m;   // r-th bead outside [s,e)
            double bx = pos[3 * b], by = pos[3 * b + 1], bz = pos[3 * b + 2];
            int p, nextp;
            native_start(nat, b, s + g, &p, &nextp);
            for (int ia = s + g; ia < e; ia += S) {
                double r0 = native_r0(nat, b, ia, &p, &nextp);
                int d = ia - b;
                if (d < 3 && d > -3) continue;   // bonded neighbours
                double dx = pos[3 * ia] - bx, dy = pos[3 * ia + 1] - by, dz = pos[3 * ia + 2] - bz;
                double ex = newp[3 * ia] - bx, ey = newp[3 * ia + 1] - by, ez = newp[3 * ia + 2] - bz;
                acc += go_pair_delta(ex * ex + ey * ey + ez * ez, dx * dx + dy * dy + dz * dz, r0, a.eps, a.sig2);
            }
        }
    } else {
        const int S = max(1, (int)blockDim.x / n);
        for (int idx = threadIdx.x; idx < n * S; idx += blockDim.x) {
            int g = idx / n, b = idx - g * n;
            bool bin = b >= s && b < e;
            double bx = pos[3 * b], by = pos[3 * b + 1], bz = pos[3 * b + 2];
            double cx = bin ? newp[3 * b] : bx, cy = bin ? newp[3 * b + 1] : by, cz = bin ? newp[3 * b + 2] : bz;
            int p, nextp;
            native_start(nat, b, s + g, &p, &nextp);
            for (int ia = s + g; ia < e; ia += S) {
                double r0 = native_r0(nat, b, ia, &p, &nextp);
                int d = ia - b;
                if (d < 3 && d > -3) continue;            // bonded neighbours and itself
                if (bin && b < ia) continue;              // pairs inside the moved range once
                double dx = pos[3 * ia] - bx, dy = pos[3 * ia + 1] - by, dz = pos[3 * ia + 2] - bz;
                double ex = newp[3 * ia] - cx, ey = newp[3 * ia + 1] - cy, ez = newp[3 * ia + 2] - cz;
                acc += go_pair_inv(recip64(ex * ex + ey * ey + ez * ez), r0, a.eps, a.sig2) -
                       go_pair_inv(recip64(dx * dx + dy * dy + dz * dz), r0, a.eps, a.sig2);
            }
        }
    }
    // bonds (cbond.c:103-104 form) with at least one end in [s,e)
    int blo = s > 0 ? s - 1 : 0, bhi = e < n ? e : n - 1;
    for (int i = blo + (int)threadIdx.x; i < bhi; i += blockDim.x) {
        int j = i + 1;
        bool im = i >= s && i < e, jm = j >= s && j < e;
        if ((!im && !jm) || (rigid && im && jm)) continue;
        double ox = pos[3 * j] - pos[3 * i], oy = pos[3 * j + 1] - pos[3 * i + 1], oz = pos[3 * j + 2] - pos[3 * i + 2];
        const double *pi = im ? newp + 3 * i : pos + 3 * i;
        const double *pj = jm ? newp + 3 * j : pos + 3 * j;
        double qx = pj[0] - pi[0], qy = pj[1] - pi[1], qz = pj[2] - pi[2];
        double ro = sqrt(ox * ox + oy * oy + oz * oz) - a.r0bond;
        double rn = sqrt(qx * qx + qy * qy + qz * qz) - a.r0bond;
        acc += a.kbond * rn * rn - a.kbond * ro * ro;
    }
    return acc;   // this thread's share
}

__device__ double range_delta(const double *pos, const double *newp, int n, int s, int e, const Natives &nat,
                              const McArgs &a, double *red, bool rigid = false) {
    return block_sum(range_delta_part(pos, newp, n, s, e, nat, a, rigid), red);
}

// The same for a one-bead displacement of bead k (mcmove_displ.c): thread b
// takes the pair (b, k), |b - k| >= 3, with the native distance nmark[b]
// (set for k's native partners, zero for every other bead), and the bonds of
// k.  This thread's share.
__device__ double displ_delta_part(const double *pos, const double *newp, int n, int k, const double *nmark,
                                   const McArgs &a) {
    double acc = 0.0;
    double ox = pos[3 * k], oy = pos[3 * k + 1], oz = pos[3 * k + 2];
    double cx = newp[3 * k], cy = newp[3 * k + 1], cz = newp[3 * k + 2];
    for (int b = threadIdx.x; b < n; b += blockDim.x) {
        int d = b - k;
        if (d < 3 && d > -3) continue;
        double r0 = nmark[b];
        double bx = pos[3 * b], by = pos[3 * b + 1], bz = pos[3 * b + 2];
        double dx = ox - bx, dy = oy - by, dz = oz - bz;
        double ex = cx - bx, ey = cy - by, ez = cz - bz;
        acc += go_pair_delta(ex * ex + ey * ey + ez * ez, dx * dx + dy * dy + dz * dz, r0, a.eps, a.sig2);
    }
    int i = k - 1 + (int)threadIdx.x;   // bonds (k-1, k) and (k, k+1), cbond.c:103-104 form
    if (threadIdx.x < 2 && i >= 0 && i + 1 < n) {
        int j = i + 1;
        const double *pi = i == k ? newp + 3 * i : pos + 3 * i;
        const double *pj = j == k ? newp + 3 * j : pos + 3 * j;
        double qx = pos[3 * j] - pos[3 * i], qy = pos[3 * j + 1] - pos[3 * i + 1], qz = pos[3 * j + 2] - pos[3 * i + 2];
        double rx = pj[0] - pi[0], ry = pj[1] - pi[1], rz = pj[2] - pi[2];
        double ro = sqrt(qx * qx + qy * qy + qz * qz) - a.r0bond;
        double rn = sqrt(rx * rx + ry * ry + rz * rz) - a.r0bond;
        acc += a.kbond * rn * rn - a.kbond * ro * ro;
    }
    return acc;
}

__device__ double full_energy(const double *pos, int n, const Natives &nat, const McArgs &a, double *red) {
    double acc = 0.0;
    const int S = max(1, (int)blockDim.x / n);
    for (int idx = threadIdx.x; idx < n * S; idx += blockDim.x) {
        int g = idx / n, b = idx - g * n;
        double bx = pos[3 * b], by = pos[3 * b + 1], bz = pos[3 * b + 2];
        int p, nextp;
        native_start(nat, b, g, &p, &nextp);
        for (int ia = g; ia + 3 <= b; ia += S) {
            double r0 = native_r0(nat, b, ia, &p, &nextp);
            double dx = pos[3 * ia] - bx, dy = pos[3 * ia + 1] - by, dz = pos[3 * ia + 2] - bz;
            acc += go_pair_inv(recip64(dx * dx + dy * dy + dz * dz), r0, a.eps, a.sig2);
        }
    }
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) {
        double dx = pos[3 * (i + 1)] - pos[3 * i], dy = pos[3 * (i + 1) + 1] - pos[3 * i + 1], dz = pos[3 * (i + 1) + 2] - pos[3 * i + 2];
        double r = sqrt(dx * dx + dy * dy + dz * dz) - a.r0bond;
        acc += a.kbond * r * r;
    }
    return block_sum(acc, red);
}

// Network weights as held in LDS by the sampling kernel.  The forward pass
// runs on waves 0-3 (one per SIMD): wave q evaluates the move (q < 2) or the
// reverse move (q >= 2), and its thread t = (q % 2) * 64 + lane takes output
// j = t / 2 of a layer over the inputs i = 32 k .. 32 k + 31, k = t % 2.
// Each layer's matrix is stored so that thread t finds its 32 weights as
// eight float4 at R[m4 * 128 + t] (words (m4 * 128 + t) * 4 + m hold input
// 32 k + 4 m4 + m); inputs past NF carry zero weight.
//   R1[4096], R2[4096], b1[HID], b2[HID], W3[HID]   (b3 cancels in g(x) - g(swap x))
static_assert(HID == 64 && NF <= 64 && MC_THREADS == 1024, "network layout assumes 64 hidden units and 16 waves");
#define NRUN (2 * 4096 + 3 * HID)
#define FEAT_SWAP 64   // offset of the reverse move's features in the feature buffer

__device__ void mlp_load_lds(const float *g, float *W) {
    const float *W1 = g, *b1 = W1 + HID * NF, *W2 = b1 + HID, *b2 = W2 + HID * HID, *W3 = b2 + HID;
    float *R1 = W, *R2 = W + 4096, *sb1 = W + 8192, *sb2 = sb1 + HID, *sW3 = sb2 + HID;
    for (int u = threadIdx.x; u < 4096; u += blockDim.x) {
        int m4 = u >> 9, t = (u >> 2) & 127, m = u & 3;
        int j = t >> 1, i = 32 * (t & 1) + 4 * m4 + m;
        R1[u] = i < NF ? W1[j * NF + i] : 0.0f;
        R2[u] = W2[j * HID + i];
    }
    for (int t = threadIdx.x; t < HID; t += blockDim.x) {
        sb1[t] = b1[t];
        sb2[t] = b2[t];
        sW3[t] = W3[t];
    }
}

// feature index of the reverse move (old <-> new), as the sampling kernel
// writes it at FEAT_SWAP
__device__ __forceinline__ int feat_swap(int i) {
    return i < 16 ? i + 16 : (i < 32 ? i - 16 : (i < 40 ? i + 8 : (i < 48 ? i - 8 : i)));
}

// 32-term dot product of thread t's weights with inputs x[0..32).  The LDS
// reads of each half are issued together (one round trip per half instead of
// one per float4: the scheduler otherwise serialises them at this kernel's
// register budget).
__device__ __forceinline__ float dot32(const float4 *R, int t, const float4 *x) {
    float acc = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float4 w[4], v[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            w[m] = R[(4 * h + m) * 128 + t];
            v[m] = x[4 * h + m];
        }
        __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);   // 8 DS reads first
        __builtin_amdgcn_sched_group_barrier(0x002, 16, 0);  // then the 16 FMAs
#pragma unroll
        for (int m = 0; m < 4; m++) {
            acc = fmaf(w[m].x, v[m].x, acc);
            acc = fmaf(w[m].y, v[m].y, acc);
            acc = fmaf(w[m].z, v[m].z, acc);
            acc = fmaf(w[m].w, v[m].w, acc);
        }
    }
    return acc;
}

// Surrogate g(x) - g(swap x) for the features x = xs[0..NF) and the reverse
// move's xs[FEAT_SWAP..FEAT_SWAP+NF) (slots NF..63 of both hold zero), with
// the weights laid out by mlp_load_lds.  Called by all 1024 threads (two
// barriers inside; waves 4-15 only wait); leaves four partial sums in pv,
// g = (pv[0] + pv[1]) + (pv[2] + pv[3]).  xs and hs 16-byte aligned; hs:
// 2*HID floats.
__device__ void mlp_block(const float *W, const float *xs, float *hs, float *pv) {
    const float4 *R1 = (const float4 *)W, *R2 = R1 + 1024;
    const float *b1 = W + 8192, *b2 = b1 + HID, *W3 = b2 + HID;
    int q = threadIdx.x >> 6, l = threadIdx.x & 63;
    int sm = q >> 1, t = ((q & 1) << 6) | l, j = t >> 1, k = t & 1;
    if (q < 4) {
        float acc = dot32(R1, t, (const float4 *)(xs + FEAT_SWAP * sm) + 8 * k);
        acc += dpp_f32<DPP_XOR1>(acc);
        if (k == 0) hs[HID * sm + j] = tanhf(acc + b1[j]);
    }
    __syncthreads();
    if (q < 4) {
        float acc = dot32(R2, t, (const float4 *)(hs + HID * sm) + 8 * k);
        acc += dpp_f32<DPP_XOR1>(acc);
        float c = k == 0 ? W3[j] * tanhf(acc + b2[j]) : 0.0f;
        c = wave_sum_f32(c);
        if (l == 0) pv[q] = sm ? -c : c;
    }
    __syncthreads();
}

// ln g at energy en in window bin b (dos_interpolate.c:107-124): on the line
// through the bin's midpoint e_begin + e_width (b + 0.5) (init.c:345) with
// the central slope of the neighbouring bins, one-sided in the first and
// last bin, in the reference's order of operations and without contraction.
// A one-bin window has no slope.
__device__ __forceinline__ double lng_interp(const double *lng, int nbins, int b, double en, double emin, double ebin) {
#pragma clang fp contract(off)
    double mid = emin + ebin * (b + 0.5);
    if (nbins < 2) return lng[b];
    if (b == 0) return lng[b] + (lng[b + 1] - lng[b]) * (en - mid) / ebin;
    if (b == nbins - 1) return lng[b] + (lng[b] - lng[b - 1]) * (en - mid) / ebin;
    return lng[b] + (lng[b + 1] - lng[b - 1]) * (en - mid) * 0.5 / ebin;
}

// u < exp(lr) for u in [0,1): decided in FP32 unless u lies within 1e-5
// (relative) of the threshold, then in FP64.  Every thread gets the same
// answer from the same inputs.
__device__ __forceinline__ bool accept_log(double lr, double u) {
    if (lr >= 0.0) return true;
    if (lr < -80.0) return u < exp(lr);   // below the FP32 range of exp
    float t = __expf((float)lr), uf = (float)u;
    if (uf < t * (1.0f - 1e-5f)) return true;
    if (uf > t * (1.0f + 1e-5f)) return false;
    return u < exp(lr);
}

// Proposals of the sampling kernel, MC_RB trials at a time: wave 0, lane l
// draws MC_RPT uniforms from its own xoshiro256** stream l of the walker
// (state srng[q * 64 + l], q < 4, in LDS) and turns them into trial l's move
// (mcmove_pivot.c:54-140: pivot about a random bead and axis, one side of the
// chain, angle u * pi * theta_max of random sign; mcmove_displ.c:45-99: one
// bead displaced by (u - 0.5) * disp_max per axis) in a table in LDS:
//   doubles tb[k * MC_RB + l]: first- and second-stage acceptance numbers,
//           signed angle, its cos and sin, displacement x, y, z
//   ints    ti[k * MC_RB + l]: type (1 pivot, 0 displacement), moved beads
//           [s, e), pivot (or displaced) bead, rotation axis
// A proposal does not depend on the walker's state, so a whole batch can be
// prepared ahead of the trials that use it.
#define MC_STREAMS 64
#define MC_RB 64
#define MC_RPT 8
#define MC_TD 8
#define MC_TI 5
__device__ __forceinline__ void draw_batch(uint64_t *srng, double *tb, int *ti, int n, const McArgs &a, const double *tc) {
    int l = threadIdx.x & 63;
    uint64_t st[4] = {srng[l], srng[64 + l], srng[128 + l], srng[192 + l]};
    double u[MC_RPT];
#pragma unroll
    for (int k = 0; k < MC_RPT; k++) u[k] = xo_uniform(st);
    srng[l] = st[0]; srng[64 + l] = st[1]; srng[128 + l] = st[2]; srng[192 + l] = st[3];
    int type, s, e, b, axis = 0;
    double th = 0.0, c = 1.0, sn = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
    if (u[0] < a.pivot_frac && n > 2) {
        int p = (int)(u[1] * n);
        int side = (int)(u[3] * 2.0);
        double ang = u[5] * MC_PI * a.theta_max;
        axis = (int)(u[2] * 3.0);
        th = (int)(u[4] * 2.0) ? -ang : ang;
        sincos_pivot(th, tc, &sn, &c);
        type = 1;
        s = side == 0 ? 0 : p + 1;
        e = side == 0 ? p : n;
        b = p;
    } else {
        int k = (int)(u[1] * n);
        type = 0;
        s = k;
        e = k + 1;
        b = k;
        d0 = (u[2] - 0.5) * a.disp_max;
        d1 = (u[3] - 0.5) * a.disp_max;
        d2 = (u[4] - 0.5) * a.disp_max;
    }
    tb[l] = u[6]; tb[MC_RB + l] = u[7]; tb[2 * MC_RB + l] = th; tb[3 * MC_RB + l] = c; tb[4 * MC_RB + l] = sn;
    tb[5 * MC_RB + l] = d0; tb[6 * MC_RB + l] = d1; tb[7 * MC_RB + l] = d2;
    ti[l] = type; ti[MC_RB + l] = s; ti[2 * MC_RB + l] = e; ti[3 * MC_RB + l] = b; ti[4 * MC_RB + l] = axis;
}

// New position of bead i under the rotation about `axis` through bead piv
// (PIVOT matrices), written out per axis (hipcc 7.2 at -O3 mis-assembled the
// z matrix when it was built in a zero-initialised 3x3 array and applied
// generically).
__device__ __forceinline__ void pivot_bead(const double *pos, int i, int piv, int axis, double c, double sn, double *o) {
    double px = pos[3 * piv], py = pos[3 * piv + 1], pz = pos[3 * piv + 2];
    double x = pos[3 * i] - px, y = pos[3 * i + 1] - py, z = pos[3 * i + 2] - pz;
    double nx, ny, nz;
    if (axis == 0) { nx = x; ny = c * y - sn * z; nz = sn * y + c * z; }
    else if (axis == 1) { nx = c * x + sn * z; ny = y; nz = c * z - sn * x; }
    else { nx = c * x + sn * y; ny = c * y - sn * x; nz = z; }
    o[0] = nx + px; o[1] = ny + py; o[2] = nz + pz;
}

// sum of the 16 waves' partial sums red[0..16), computed by one wave in a
// fixed order; wave-uniform result
__device__ __forceinline__ double sum_waves(const double *red) {
    int l = threadIdx.x & 63;
    return wave_sum_f64(l < MC_WAVES ? red[l] : 0.0);
}

// One workgroup per walker runs all its trials.  The work that is the same
// for the whole walker (drawing proposals, Wang-Landau bookkeeping,
// acceptance, copying accepted coordinates, logging) runs once, in wave 0;
// all 16 waves read the trial's move from the proposal table and compute the
// moved coordinates, the move features, the network and the exact energy
// change.  Barriers per trial: moved coordinates and features ready; (delayed
// acceptance: between the layers, after the network, after the first-stage
// decision); exact partial sums ready; bookkeeping done.
template <bool DA>
__global__ void __launch_bounds__(MC_THREADS)
k_mc_run(McArgs a, int ntrials, double lnf, int do_log, double *gpos, const int *gn, const int *gnoff,
         const int *gnpart, const double *gnr0, int ncap, int max_nnz, const double *gemin, double *gecur, double *glng, long long *ghist, uint64_t *grng, long long *gstats,
         const float *gmlp, float *logx, double *logde, int *logn, int logcap, double escale) {
    extern __shared__ double smem[];
    int w = blockIdx.x;
    int n = gn[w];
    if (n <= 0) return;
    double *pos = smem;                       // 3*maxn
    double *newp = pos + 3 * a.maxn;          // 3*maxn
    double *lng = newp + 3 * a.maxn;          // nbins
    long long *hist = (long long *)(lng + a.nbins);
    double *red = (double *)(hist + a.nbins); // MC_WAVES
    // (offsets are aligned by arithmetic on the LDS base pointer: a round
    // trip through an integer would make every access a flat one)
    size_t wofs = (((char *)(red + MC_WAVES) - (char *)smem) + 15) & ~(size_t)15;
    float *W = (float *)((char *)smem + wofs);   // NRUN (16-byte aligned)
    float *feat = W + (DA ? NRUN : 0);        // 2*FEAT_SWAP: move, reverse move
    float *hs = feat + 2 * FEAT_SWAP;         // 2*HID
    float *pv = hs + 2 * HID;                 // MC_WAVES
    size_t rofs = (((char *)(pv + MC_WAVES) - (char *)smem) + 7) & ~(size_t)7;
    double *sr0 = (double *)((char *)smem + rofs);   // max_nnz
    int *soff = (int *)(sr0 + max_nnz);       // maxn + 1
    int *spart = soff + a.maxn + 1;           // max_nnz
    size_t uofs = (((char *)(spart + max_nnz) - (char *)smem) + 15) & ~(size_t)15;
    double *tb = (double *)((char *)smem + uofs);   // MC_TD * MC_RB: proposal table (doubles)
    uint64_t *srng = (uint64_t *)(tb + MC_TD * MC_RB);   // 4 * MC_STREAMS: random streams
    float4 *posf = (float4 *)(srng + 4 * MC_STREAMS);      // maxn: FP32 copy of pos for the features
    double *tcoef = (double *)(posf + a.maxn);   // 17: sin/cos Taylor coefficients
    int *ti = (int *)(tcoef + 17);            // MC_TI * MC_RB: proposal table (ints)
    int *pass = ti + MC_TI * MC_RB;           // first-stage decision
    size_t mofs = (((char *)(pass + 1) - (char *)smem) + 7) & ~(size_t)7;
    double *nmark = (double *)((char *)smem + mofs);   // maxn: native distance to the displaced bead

    for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) pos[i] = gpos[(size_t)w * 3 * a.maxn + i];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double *g = gpos + (size_t)w * 3 * a.maxn + 3 * i;
        posf[i] = make_float4((float)g[0], (float)g[1], (float)g[2], 0.0f);
    }
    for (int i = threadIdx.x; i <= n; i += blockDim.x) soff[i] = gnoff[(size_t)w * (a.maxn + 1) + i];
    {
        int nnz = gnoff[(size_t)w * (a.maxn + 1) + n];
        for (int i = threadIdx.x; i < nnz; i += blockDim.x) {
            spart[i] = gnpart[(size_t)w * ncap + i];
            sr0[i] = gnr0[(size_t)w * ncap + i];
        }
    }
    const Natives nat = {soff, spart, sr0};
    for (int i = threadIdx.x; i < a.nbins; i += blockDim.x) {
        lng[i] = glng[(size_t)w * a.nbins + i];
        hist[i] = ghist[(size_t)w * a.nbins + i];
    }
    if (DA) mlp_load_lds(gmlp, W);
    for (int i = threadIdx.x; i < 2 * FEAT_SWAP; i += blockDim.x) feat[i] = 0.0f;   // slots past NF stay zero
    if (threadIdx.x < 17) tcoef[threadIdx.x] = kSinCosTaylor[threadIdx.x];
    for (int i = threadIdx.x; i < n; i += blockDim.x) nmark[i] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool w0 = wv == 0;
    // wave 0's walker state
    const double emin = gemin[w];
    double ecur = gecur[w];
    // Wang-Landau bin of an energy: floor((E - emin) / ebin), with the
    // division done as a multiplication by the reciprocal (the bin edges move
    // by at most an ulp of the quotient; every bin here uses the same form)
    const double inv_ebin = 1.0 / a.ebin;
    auto bin_of = [&](double en) { return (int)floor((en - emin) * inv_ebin); };
    auto lg = [&](int b, double en) { return lng_interp(lng, a.nbins, b, en, emin, a.ebin); };
    int bx = bin_of(ecur);   // bin of the current energy
    double sur = 0.0, a1f = 0.0;   // surrogate dE, log first-stage probability of the trial
    int st_acc = 0, st_s1 = 0, st_ex = 0, st_out = 0, st_piv = 0, st_pacc = 0;   // this launch
    if (w0) {
        const uint64_t *g = grng + ((size_t)w * MC_STREAMS + lane) * 4;
        for (int q = 0; q < 4; q++) srng[q * 64 + lane] = g[q];
        draw_batch(srng, tb, ti, n, a, tcoef);
    }
    __syncthreads();

    // Wave 0: the Wang-Landau update of the bin the trial ended in
    // (pivot.c:236-253), and the next batch of proposals when this one is used up.
    auto finish = [&](int trial) {
        if (bx >= 0 && bx < a.nbins && lane == 0) {   // LDS atomics: nothing waits for them
            atomicAdd(&lng[bx], lnf);
            atomicAdd((unsigned long long *)&hist[bx], 1ull);
        }
        if ((trial + 1) % MC_RB == 0 && trial + 1 < ntrials) draw_batch(srng, tb, ti, n, a, tcoef);
    };

    for (int trial = 0; trial < ntrials; trial++) {
        const int tl = trial % MC_RB;   // this trial's column of the proposal table
        const int type = __builtin_amdgcn_readfirstlane(ti[tl]);
        const int s = __builtin_amdgcn_readfirstlane(ti[MC_RB + tl]), e = __builtin_amdgcn_readfirstlane(ti[2 * MC_RB + tl]);
        const int piv = __builtin_amdgcn_readfirstlane(ti[3 * MC_RB + tl]), axis = __builtin_amdgcn_readfirstlane(ti[4 * MC_RB + tl]);
        const double c = tb[3 * MC_RB + tl], sn = tb[4 * MC_RB + tl];
        // ---- moved coordinates
        if (type == 1) {
            for (int i = s + threadIdx.x; i < e; i += blockDim.x) pivot_bead(pos, i, piv, axis, c, sn, newp + 3 * i);
        } else {
            if (threadIdx.x == 0) {
                newp[3 * s] = pos[3 * s] + tb[5 * MC_RB + tl];
                newp[3 * s + 1] = pos[3 * s + 1] + tb[6 * MC_RB + tl];
                newp[3 * s + 2] = pos[3 * s + 2] + tb[7 * MC_RB + tl];
            }
            // native distances to the displaced bead (wave 0 clears them again)
            int o = soff[s], dg = soff[s + 1] - o;
            if ((int)threadIdx.x < dg) nmark[spart[o + threadIdx.x]] = sr0[o + threadIdx.x];
        }
        const bool move = e > s;   // empty moved part: nothing changes, a rejected trial
        // ---- features of pivot moves: local distances around the pivot, and
        // the closest approach of sampled moved beads to the fixed part (the
        // waves rotate the beads they need themselves)
        if (type == 1 && move && (DA || do_log)) {
            int sg = (s == 0) ? -1 : 1;     // moved side direction from the pivot
            float sig2f = (float)a.sig2;
            if (wv < 8) {
                // wave k < 8 takes sampled bead k: lanes scan the fixed beads,
                // the minima are reduced in the wave
                int bk = piv + sg * (2 << wv);
                float mo = 3.0e38f, mn = 3.0e38f;
                if (bk >= s && bk < e) {
                    float4 b0 = posf[bk];
                    double nb[3];
                    pivot_bead(pos, bk, piv, axis, c, sn, nb);
                    float bx1 = (float)nb[0], by1 = (float)nb[1], bz1 = (float)nb[2];
                    // the fixed part of a pivot move is one side, [e, n) or [0, s)
                    int flo = s == 0 ? e : 0, fhi = s == 0 ? n : s;
#pragma unroll 4
                    for (int f = flo + lane; f < fhi; f += 64) {
                        int d = f - bk;
                        float4 pf = posf[f];
                        float ox = b0.x - pf.x, oy = b0.y - pf.y, oz = b0.z - pf.z;
                        float nx = bx1 - pf.x, ny = by1 - pf.y, nz = bz1 - pf.z;
                        bool far = d >= 3 || d <= -3;
                        mo = far ? fminf(mo, ox * ox + oy * oy + oz * oz) : mo;
                        mn = far ? fminf(mn, nx * nx + ny * ny + nz * nz) : mn;
                    }
                }
                mo = wave_min_f32(mo);
                mn = wave_min_f32(mn);
                if (lane == 0) {
                    float fo = bf16_round(mo < 1e37f ? feat_term(mo, sig2f) : 0.0f);
                    float fnw = bf16_round(mn < 1e37f ? feat_term(mn, sig2f) : 0.0f);
                    feat[32 + wv] = fo;
                    feat[40 + wv] = fnw;
                    feat[FEAT_SWAP + 32 + wv] = fnw;
                    feat[FEAT_SWAP + 40 + wv] = fo;
                }
            } else if (wv == 8) {
                // local distances of the 4 moved x 4 fixed beads nearest the
                // pivot, and the move's scalars
                if (lane < 16) {
                    int i = lane / 4 + 1, j = lane % 4 + 1;
                    int mb = piv + sg * i, fb = piv - sg * j;
                    float v = 0.0f, v2 = 0.0f;
                    if (mb >= 0 && mb < n && fb >= 0 && fb < n) {
                        float4 pm = posf[mb], pf = posf[fb];
                        double nb[3];
                        pivot_bead(pos, mb, piv, axis, c, sn, nb);
                        float ox = pm.x - pf.x, oy = pm.y - pf.y, oz = pm.z - pf.z;
                        float nx = (float)nb[0] - pf.x, ny = (float)nb[1] - pf.y, nz = (float)nb[2] - pf.z;
                        v = bf16_round(feat_term(ox * ox + oy * oy + oz * oz, sig2f));
                        v2 = bf16_round(feat_term(nx * nx + ny * ny + nz * nz, sig2f));
                    }
                    feat[lane] = v;
                    feat[16 + lane] = v2;
                    feat[FEAT_SWAP + lane] = v2;
                    feat[FEAT_SWAP + 16 + lane] = v;
                } else if (lane < 20) {
                    int q = lane - 16;
                    float v = bf16_round(q == 0 ? 1.0f : q == 1 ? (float)(fabs(tb[2 * MC_RB + tl]) * (1.0 / MC_PI))
                                                 : q == 2 ? (float)(e - s) / (float)n : 0.0f);
                    feat[48 + q] = v;
                    feat[FEAT_SWAP + 48 + q] = v;
                }
            }
        }
        __syncthreads();   // moved coordinates and features are visible
        if (w0) st_piv += type;
        if (DA && type == 1 && move) {
            // first stage: surrogate dE of the move (antisymmetric network),
            // a1 = min(1, g(E_x) / g(E_x + sur)) as log a1, g interpolated
            mlp_block(W, feat, hs, pv);
            if (w0) {
                float ex = __expf((pv[0] + pv[1]) + (pv[2] + pv[3]));
                float so = 0.5f * (ex - __builtin_amdgcn_rcpf(ex));   // sinh of the network output
                sur = escale * (double)so;
                int by = bin_of(ecur + sur);
                a1f = (bx >= 0 && bx < a.nbins && by >= 0 && by < a.nbins) ? fmin(0.0, lg(bx, ecur) - lg(by, ecur + sur))
                                                                          : -INFINITY;
                bool ok = accept_log(a1f, tb[tl]);
                if (lane == 0) *pass = ok ? 1 : 0;
                if (!ok) {
                    st_s1++;
                    finish(trial);
                }
            }
            __syncthreads();   // first-stage decision (and, after a rejection, the bookkeeping) visible
            if (!__builtin_amdgcn_readfirstlane(*pass)) continue;
        }
        // ---- exact range energy change (FP64), partial sums per wave
        if (move) {
            double part = type == 1 ? range_delta_part(pos, newp, n, s, e, nat, a, true)
                                    : displ_delta_part(pos, newp, n, s, nmark, a);
            part = wave_sum_f64(part);
            if (lane == 0) red[wv] = part;
            __syncthreads();
        }
        if (w0) {
            bool acc = false;
            if (move) {
                double de = sum_waves(red);
                st_ex++;
                int by = bin_of(ecur + de);
                if (by < 0 || by >= a.nbins || bx < 0 || bx >= a.nbins) {
                    st_out++;
                } else {
                    // pivot.c:236-253: arg = ln g(old) - ln g(new), both
                    // interpolated at the energies (dos_interp)
                    double lratio = lg(bx, ecur) - lg(by, ecur + de);
                    if (DA && type == 1) {
                        // second stage: pi(y) a1(y,x) / (pi(x) a1(x,y)), a1(y,x) from
                        // the surrogate of the reverse move, -sur (all as logarithms)
                        int bz = bin_of(ecur + de - sur);
                        double la1r = -INFINITY;
                        if (bz >= 0 && bz < a.nbins) la1r = fmin(0.0, lg(by, ecur + de) - lg(bz, ecur + de - sur));
                        lratio = lratio + la1r - a1f;
                    }
                    acc = accept_log(lratio, tb[MC_RB + tl]);
                }
                if (do_log && type == 1) {
                    int slot = 0;
                    if (lane == 0) slot = atomicAdd(logn, 1);
                    slot = __builtin_amdgcn_readfirstlane(slot);
                    if (slot < logcap) {
                        for (int q = lane; q < NF; q += 64) logx[(size_t)slot * NF + q] = feat[q];
                        if (lane == 0) logde[slot] = de;
                    }
                }
                if (acc) {
                    // moved beads into pos: two per lane per pass, loads first
                    for (int i0 = s + lane; i0 < e; i0 += 128) {
                        double v[2][3];
#pragma unroll
                        for (int u = 0; u < 2; u++) {
                            int i = i0 + 64 * u;
                            if (i < e) { v[u][0] = newp[3 * i]; v[u][1] = newp[3 * i + 1]; v[u][2] = newp[3 * i + 2]; }
                        }
#pragma unroll
                        for (int u = 0; u < 2; u++) {
                            int i = i0 + 64 * u;
                            if (i < e) {
                                pos[3 * i] = v[u][0]; pos[3 * i + 1] = v[u][1]; pos[3 * i + 2] = v[u][2];
                                posf[i] = make_float4((float)v[u][0], (float)v[u][1], (float)v[u][2], 0.0f);
                            }
                        }
                    }
                    ecur += de;
                    bx = by;   // bin_of(ecur) of the new ecur
                    st_acc++;
                    st_pacc += type;
                }
                if (type == 0) {
                    int o = soff[s], dg = soff[s + 1] - o;
                    for (int q = lane; q < dg; q += 64) nmark[spart[o + q]] = 0.0;
                }
            }
            finish(trial);
        }
        __syncthreads();   // bookkeeping (accepted coordinates, next proposals) visible
    }
    for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) gpos[(size_t)w * 3 * a.maxn + i] = pos[i];
    for (int i = threadIdx.x; i < a.nbins; i += blockDim.x) {
        glng[(size_t)w * a.nbins + i] = lng[i];
        ghist[(size_t)w * a.nbins + i] = hist[i];
    }
    if (w0) {
        uint64_t *g = grng + ((size_t)w * MC_STREAMS + lane) * 4;
        for (int q = 0; q < 4; q++) g[q] = srng[q * 64 + lane];
    }
    if (threadIdx.x == 0) {
        gecur[w] = ecur;
        long long *st = gstats + (size_t)w * TTC_MC_NSTAT;
        st[0] += ntrials; st[1] += st_acc; st[2] += st_s1; st[3] += st_ex; st[4] += st_out;
        st[5] += st_piv; st[6] += st_pacc;
    }
}

__global__ void __launch_bounds__(MC_THREADS)
k_mc_energy(McArgs a, const double *gpos, const int *gn, const int *gnoff, const int *gnpart, const double *gnr0,
            int ncap, double *eout) {
    extern __shared__ double smem[];
    int w = blockIdx.x;
    int n = gn[w];
    double *pos = smem, *red = pos + 3 * a.maxn;
    for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) pos[i] = gpos[(size_t)w * 3 * a.maxn + i];
    __syncthreads();
    const Natives nat = {gnoff + (size_t)w * (a.maxn + 1), gnpart + (size_t)w * ncap, gnr0 + (size_t)w * ncap};
    double e = n > 0 ? full_energy(pos, n, nat, a, red) : 0.0;
    if (threadIdx.x == 0) eout[w] = e;
}

__global__ void __launch_bounds__(MC_THREADS)
k_mc_delta(McArgs a, int w, int s, int e, const double *gpos, const int *gn, const int *gnoff, const int *gnpart,
           const double *gnr0, int ncap, const double *newrange, double *out) {
    extern __shared__ double smem[];
    int n = gn[w];
    double *pos = smem, *newp = pos + 3 * a.maxn, *red = newp + 3 * a.maxn;
    for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) pos[i] = gpos[(size_t)w * 3 * a.maxn + i];
    for (int i = threadIdx.x; i < 3 * (e - s); i += blockDim.x) newp[3 * s + i] = newrange[i];
    __syncthreads();
    const Natives nat = {gnoff + (size_t)w * (a.maxn + 1), gnpart + (size_t)w * ncap, gnr0 + (size_t)w * ncap};
    double d = range_delta(pos, newp, n, s, e, nat, a, red);
    if (threadIdx.x == 0) *out = d;
}

// ---- network training / evaluation (one workgroup, 256 threads) -----------

#define TR_THREADS 256
#define TR_B 32   // samples per minibatch step (each evaluated twice: x and swapped x)

// forward of R rows of features X (R x NF, already bf16) -> H1, H2 (R x HID), out (R)
__device__ void mlp_fwd_rows(const float *P, const float *X, int R, float *H1, float *H2, float *out) {
    const float *W1 = P, *b1 = W1 + HID * NF, *W2 = b1 + HID, *b2 = W2 + HID * HID, *W3 = b2 + HID, *b3 = W3 + HID;
    for (int t = threadIdx.x; t < R * HID; t += blockDim.x) {
        int r = t / HID, j = t % HID;
        float s = b1[j];
        for (int i = 0; i < NF; i++) s += W1[j * NF + i] * X[r * NF + i];
        H1[t] = tanhf(s);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * HID; t += blockDim.x) {
        int r = t / HID, j = t % HID;
        float s = b2[j];
        for (int i = 0; i < HID; i++) s += W2[j * HID + i] * H1[r * HID + i];
        H2[t] = tanhf(s);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        float s = b3[0];
        for (int j = 0; j < HID; j++) s += W3[j] * H2[r * HID + j];
        out[r] = s;
    }
    __syncthreads();
}

// gradient of sum_r dout[r] * f(X_r) w.r.t. P, written (not added) to G
__device__ void mlp_bwd_rows(const float *P, const float *X, int R, const float *H1, const float *H2,
                             const float *dout, float *D2, float *D1, float *G) {
    const float *W2 = P + HID * NF + HID, *W3 = W2 + HID * HID + HID;
    float *gW1 = G, *gb1 = gW1 + HID * NF, *gW2 = gb1 + HID, *gb2 = gW2 + HID * HID, *gW3 = gb2 + HID, *gb3 = gW3 + HID;
    for (int t = threadIdx.x; t < R * HID; t += blockDim.x) {
        int r = t / HID, j = t % HID;
        float h = H2[t];
        D2[t] = dout[r] * W3[j] * (1.0f - h * h);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * HID; t += blockDim.x) {
        int r = t / HID, i = t % HID;
        float s = 0.0f;
        for (int j = 0; j < HID; j++) s += W2[j * HID + i] * D2[r * HID + j];
        float h = H1[t];
        D1[t] = s * (1.0f - h * h);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < HID; t += blockDim.x) {
        float s3 = 0.0f, sb2 = 0.0f, sb1 = 0.0f;
        for (int r = 0; r < R; r++) {
            s3 += dout[r] * H2[r * HID + t];
            sb2 += D2[r * HID + t];
            sb1 += D1[r * HID + t];
        }
        gW3[t] = s3;
        gb2[t] = sb2;
        gb1[t] = sb1;
    }
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int r = 0; r < R; r++) s += dout[r];
        gb3[0] = s;
    }
    for (int t = threadIdx.x; t < HID * HID; t += blockDim.x) {
        int j = t / HID, i = t % HID;
        float s = 0.0f;
        for (int r = 0; r < R; r++) s += D2[r * HID + j] * H1[r * HID + i];
        gW2[t] = s;
    }
    for (int t = threadIdx.x; t < HID * NF; t += blockDim.x) {
        int j = t / NF, i = t % NF;
        float s = 0.0f;
        for (int r = 0; r < R; r++) s += D1[r * HID + j] * X[r * NF + i];
        gW1[t] = s;
    }
    __syncthreads();
}

// plain (non-antisymmetric) loss and gradient over nsamp rows, chunked
__global__ void __launch_bounds__(TR_THREADS)
k_mlp_loss_grad(const float *P, int nsamp, const float *X, const float *Y, double *loss, float *grad) {
    extern __shared__ float sm[];
    float *Xs = sm;                       // TR_B*NF
    float *H1 = Xs + TR_B * NF, *H2 = H1 + TR_B * HID, *D2 = H2 + TR_B * HID, *D1 = D2 + TR_B * HID;
    float *out = D1 + TR_B * HID, *dout = out + TR_B, *G = dout + TR_B;
    __shared__ double lsum;
    if (threadIdx.x == 0) lsum = 0.0;
    for (int t = threadIdx.x; t < NPAR; t += blockDim.x) grad[t] = 0.0f;
    __syncthreads();
    for (int c0 = 0; c0 < nsamp; c0 += TR_B) {
        int R = min(TR_B, nsamp - c0);
        for (int t = threadIdx.x; t < R * NF; t += blockDim.x) Xs[t] = bf16_round(X[(size_t)c0 * NF + t]);
        __syncthreads();
        mlp_fwd_rows(P, Xs, R, H1, H2, out);
        if (threadIdx.x == 0) {
            for (int r = 0; r < R; r++) {
                float d = out[r] - Y[c0 + r];
                lsum += 0.5 * (double)d * d;
                dout[r] = d / (float)nsamp;
            }
        }
        __syncthreads();
        mlp_bwd_rows(P, Xs, R, H1, H2, dout, D2, D1, G);
        for (int t = threadIdx.x; t < NPAR; t += blockDim.x) grad[t] += G[t];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = lsum / nsamp;
}

__global__ void __launch_bounds__(TR_THREADS)
k_mlp_forward(const float *P, int nsamp, const float *X, float *Y) {
    extern __shared__ float sm[];
    float *Xs = sm, *H1 = Xs + TR_B * NF, *H2 = H1 + TR_B * HID, *out = H2 + TR_B * HID;
    for (int c0 = 0; c0 < nsamp; c0 += TR_B) {
        int R = min(TR_B, nsamp - c0);
        for (int t = threadIdx.x; t < R * NF; t += blockDim.x) Xs[t] = bf16_round(X[(size_t)c0 * NF + t]);
        __syncthreads();
        mlp_fwd_rows(P, Xs, R, H1, H2, out);
        for (int r = threadIdx.x; r < R; r += blockDim.x) Y[c0 + r] = out[r];
        __syncthreads();
    }
}

// Adam over the logged moves; antisymmetric output g(x) - g(swap x)
__global__ void __launch_bounds__(TR_THREADS)
// targets asinh(clip(dE)/escale): energy changes beyond declip (twice the
// Wang-Landau window) are all rejected alike and are clipped so that they do
// not dominate the loss
k_mlp_train(float *P, float *adam, int t0, int nsamp, const float *X, const double *DE, int nepoch, float lr,
            double escale, double declip, uint64_t seed, double *loss_out) {
    extern __shared__ float sm[];
    const int R = 2 * TR_B;
    float *Xs = sm;                        // R*NF
    float *H1 = Xs + R * NF, *H2 = H1 + R * HID, *D2 = H2 + R * HID, *D1 = D2 + R * HID;
    float *out = D1 + R * HID, *dout = out + R, *G = dout + R;
    __shared__ double lsum;
    __shared__ int idx[TR_B];
    uint64_t s4[4] = {seed ^ 0x9E3779B97F4A7C15ull, seed * 3 + 1, seed * 7 + 5, seed * 11 + 13};
    int t = t0;
    double last = 0.0;
    for (int ep = 0; ep < nepoch; ep++) {
        if (threadIdx.x == 0) lsum = 0.0;
        int nstep = nsamp / TR_B;
        for (int st = 0; st < nstep; st++) {
            if (threadIdx.x == 0)
                for (int r = 0; r < TR_B; r++) idx[r] = (int)(xo_uniform(s4) * nsamp) % nsamp;
            __syncthreads();
            for (int q = threadIdx.x; q < TR_B * NF; q += blockDim.x) {
                int r = q / NF, i = q % NF;
                const float *x = X + (size_t)idx[r] * NF;
                // swapped row: old <-> new local terms
                int si = feat_swap(i);
                Xs[r * NF + i] = bf16_round(x[i]);
                Xs[(TR_B + r) * NF + i] = bf16_round(x[si]);
            }
            __syncthreads();
            mlp_fwd_rows(P, Xs, R, H1, H2, out);
            if (threadIdx.x == 0) {
                for (int r = 0; r < TR_B; r++) {
                    double y = asinh(fmax(-declip, fmin(declip, DE[idx[r]])) / escale);
                    float d = (out[r] - out[TR_B + r]) - (float)y;
                    lsum += 0.5 * (double)d * d;
                    dout[r] = d / (float)TR_B;
                    dout[TR_B + r] = -d / (float)TR_B;
                }
            }
            __syncthreads();
            mlp_bwd_rows(P, Xs, R, H1, H2, dout, D2, D1, G);
            t++;
            float b1c = 1.0f - powf(0.9f, (float)t), b2c = 1.0f - powf(0.999f, (float)t);
            for (int q = threadIdx.x; q < NPAR; q += blockDim.x) {
                float g = G[q];
                float m = adam[q] = 0.9f * adam[q] + 0.1f * g;
                float v = adam[NPAR + q] = 0.999f * adam[NPAR + q] + 0.001f * g * g;
                P[q] -= lr * (m / b1c) / (sqrtf(v / b2c) + 1e-8f);
            }
            __syncthreads();
        }
        last = nstep > 0 ? lsum / (nstep * TR_B) : 0.0;
    }
    if (threadIdx.x == 0) *loss_out = last;
}

// ---------------------------------------------------------------- host

static McArgs mc_args(const ttc_mc *m) {
    McArgs a;
    a.nwalk = m->p.nwalk; a.maxn = m->p.maxn; a.nbins = m->p.nbins;
    a.ebin = m->p.ebin; a.kbond = m->p.kbond; a.r0bond = m->p.r0bond; a.eps = m->p.eps;
    a.sig2 = m->p.sig_rep * m->p.sig_rep; a.pivot_frac = m->p.pivot_frac; a.theta_max = m->p.theta_max;
    a.disp_max = m->p.disp_max;
    return a;
}

static size_t run_shm(const ttc_mc *m, bool da) {
    size_t d = (size_t)(6 * m->p.maxn + m->p.nbins) * sizeof(double) + (size_t)m->p.nbins * sizeof(long long) +
               MC_WAVES * sizeof(double);
    size_t f = 16 + (size_t)((da ? NRUN : 0) + 2 * FEAT_SWAP + 2 * HID + MC_WAVES) * sizeof(float);
    // native contacts: r0 (aligned up to 8 bytes), offsets, partners
    size_t nat = 8 + (size_t)m->max_nnz * (sizeof(double) + sizeof(int)) + (size_t)(m->p.maxn + 1) * sizeof(int);
    // random-number batches (aligned up to 16 bytes) and the FP32 bead copy
    size_t rn = 16 + (size_t)MC_TD * MC_RB * sizeof(double) + (size_t)4 * MC_STREAMS * sizeof(uint64_t) +
                (size_t)m->p.maxn * 4 * sizeof(float) + 17 * sizeof(double) + (size_t)(MC_TI * MC_RB + 1) * sizeof(int) +
                8 + (size_t)m->p.maxn * sizeof(double);
    return d + f + nat + rn;
}

// LDS of the sampling kernel on gfx950 (160 KB per workgroup)
#define MC_MAX_LDS (160 * 1024)

static size_t train_shm() {
    const int R = 2 * TR_B;
    return (size_t)(R * NF + 4 * R * HID + 2 * R + NPAR) * sizeof(float);
}

static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

#define MCHECK(cond, code, msg) \
    do {                        \
        if (!(cond)) {          \
            ttc_set_error(msg); \
            return code;        \
        }                       \
    } while (0)

// ---- scratch of the entry point that allocates nothing -------------------
// ttc_mc_delta_energy_noalloc passes the moved beads' new coordinates and
// the energy change through this array instead of per-call allocations:
// [0] the result, [NOALLOC_STAGE ...] up to TTC_MC_NOALLOC_MAXBEADS x 3
// coordinates.  Every GPU holds its own copy until the process ends.  Host
// code never passes the array's name to a kernel or a copy (in host code it
// is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns, taken once per device by ttc_mc_create with
// that device current.
#define NOALLOC_STAGE 32   // doubles before the staging area (256 B)
#define NOALLOC_MAXDEV 64
static __device__ __align__(256) double scratch[NOALLOC_STAGE + 3 * TTC_MC_NOALLOC_MAXBEADS];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAXDEV] = {};

static int scratch_setup(int dev) {
    MCHECK(dev >= 0 && dev < NOALLOC_MAXDEV, TTC_ERR_ARG, "ttc_mc_create: device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return TTC_OK;
    double *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    MCHECK(ptr, TTC_ERR_HIP, "ttc_mc_create: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
    return TTC_OK;
}

extern "C" int ttc_mlp_nparams(void) { return NPAR; }

extern "C" int ttc_mc_create(const ttc_mc_params *p, ttc_mc **out) {
    MCHECK(p && out, TTC_ERR_ARG, "ttc_mc_create: null argument");
    MCHECK(p->nwalk > 0 && p->maxn >= 4 && p->maxn <= 1024 && p->nbins > 0 && p->nbins <= 4096 && p->ebin > 0,
           TTC_ERR_ARG, "ttc_mc_create: bad sizes");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        ttc_set_error("ttc_mc_create: no HIP device");
        return TTC_ERR_HIP;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    int rs0 = scratch_setup(dev);
    if (rs0) return rs0;
    ttc_mc *m = new ttc_mc();
    m->p = *p;
    m->dev = dev;
    size_t W = p->nwalk, N = p->maxn;
    HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&m->pos, sizeof(double) * W * N * 3));
    HIPCHK(hipMalloc(&m->nbead, sizeof(int) * W));
    m->ncap = (int)N * TTC_MC_MAXNAT;
    HIPCHK(hipMalloc(&m->noff, sizeof(int) * W * (N + 1)));
    HIPCHK(hipMalloc(&m->npart, sizeof(int) * W * m->ncap));
    HIPCHK(hipMalloc(&m->nr0, sizeof(double) * W * m->ncap));
    HIPCHK(hipMalloc(&m->emin, sizeof(double) * W));
    HIPCHK(hipMalloc(&m->ecur, sizeof(double) * W));
    HIPCHK(hipMalloc(&m->lng, sizeof(double) * W * p->nbins));
    HIPCHK(hipMalloc(&m->hist, sizeof(long long) * W * p->nbins));
    HIPCHK(hipMalloc(&m->rng, sizeof(uint64_t) * W * MC_STREAMS * 4));
    HIPCHK(hipMalloc(&m->stats, sizeof(long long) * W * TTC_MC_NSTAT));
    HIPCHK(hipMalloc(&m->mlp, sizeof(float) * NPAR));
    HIPCHK(hipMalloc(&m->adam, sizeof(float) * 2 * NPAR));
    HIPCHK(hipMemset(m->nbead, 0, sizeof(int) * W));
    HIPCHK(hipMemset(m->noff, 0, sizeof(int) * W * (N + 1)));
    HIPCHK(hipMemset(m->lng, 0, sizeof(double) * W * p->nbins));
    HIPCHK(hipMemset(m->hist, 0, sizeof(long long) * W * p->nbins));
    HIPCHK(hipMemset(m->stats, 0, sizeof(long long) * W * TTC_MC_NSTAT));
    HIPCHK(hipMemset(m->adam, 0, sizeof(float) * 2 * NPAR));
    m->hn.assign(W, 0);
    m->chain_sig.assign(W, 0);
    // network: small uniform init (seeded)
    std::vector<float> w0(NPAR);
    uint64_t sx = p->seed ^ 0x5DEECE66Dull;
    for (int i = 0; i < NPAR; i++) w0[i] = (float)(((splitmix(sx) >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 0.3);
    HIPCHK(hipMemcpy(m->mlp, w0.data(), sizeof(float) * NPAR, hipMemcpyHostToDevice));
    std::vector<uint64_t> rs((size_t)W * MC_STREAMS * 4);
    uint64_t sd = p->seed;
    for (auto &v : rs) v = splitmix(sd);
    HIPCHK(hipMemcpy(m->rng, rs.data(), sizeof(uint64_t) * rs.size(), hipMemcpyHostToDevice));
    m->logcap = 1 << 20;
    HIPCHK(hipMalloc(&m->logx, sizeof(float) * (size_t)m->logcap * NF));
    HIPCHK(hipMalloc(&m->logde, sizeof(double) * m->logcap));
    HIPCHK(hipMalloc(&m->logn, sizeof(int)));
    HIPCHK(hipMemset(m->logn, 0, sizeof(int)));
    const void *tfns[] = {(const void *)k_mlp_train, (const void *)k_mlp_loss_grad, (const void *)k_mlp_forward};
    for (const void *f : tfns)
        HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)train_shm()));
    *out = m;
    return TTC_OK;
}

extern "C" int ttc_mc_destroy(ttc_mc *m) {
    if (!m) return TTC_OK;
    hipStreamSynchronize(m->stream);
    void *b[] = {m->pos, m->nbead, m->noff, m->npart, m->nr0, m->emin, m->ecur, m->lng, m->hist, m->rng, m->stats, m->mlp, m->adam,
                 m->logx, m->logde, m->logn, m->ck_pos, m->ck_ecur, m->ck_lng, m->ck_hist, m->ck_rng};
    for (void *x : b)
        if (x) hipFree(x);
    hipStreamDestroy(m->stream);
    delete m;
    return TTC_OK;
}

extern "C" int ttc_mc_set_walker(ttc_mc *m, int w, int n, const double *pos, int nnat, const int *na, const int *nb,
                                 const double *nr0, double emin) {
    MCHECK(m && pos && w >= 0 && w < m->p.nwalk && n >= 4 && n <= m->p.maxn, TTC_ERR_ARG,
           "ttc_mc_set_walker: bad argument");
    size_t N = m->p.maxn;
    // native contacts as CSR over beads, both directions, partners ascending
    std::vector<std::vector<std::pair<int, double>>> adj(n);
    for (int c = 0; c < nnat; c++) {
        int a = na[c], b = nb[c];
        MCHECK(a >= 0 && a < n && b >= 0 && b < n && abs(a - b) >= 3 && nr0[c] > 0, TTC_ERR_ARG,
               "ttc_mc_set_walker: bad native contact");
        adj[a].push_back({b, nr0[c]});
        adj[b].push_back({a, nr0[c]});
    }
    std::vector<int> off(n + 1, 0), part;
    std::vector<double> rr;
    for (int i = 0; i < n; i++) {
        std::sort(adj[i].begin(), adj[i].end());
        for (size_t k = 1; k < adj[i].size(); k++)
            MCHECK(adj[i][k].first != adj[i][k - 1].first, TTC_ERR_ARG, "ttc_mc_set_walker: repeated native contact");
        MCHECK((int)adj[i].size() <= TTC_MC_MAXNAT, TTC_ERR_ARG,
               "ttc_mc_set_walker: more than TTC_MC_MAXNAT native contacts on one bead");
        for (auto &q : adj[i]) { part.push_back(q.first); rr.push_back(q.second); }
        off[i + 1] = (int)part.size();
    }
    HIPCHK(hipMemcpy(m->noff + (size_t)w * (N + 1), off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (!part.empty()) {
        HIPCHK(hipMemcpy(m->npart + (size_t)w * m->ncap, part.data(), sizeof(int) * part.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->nr0 + (size_t)w * m->ncap, rr.data(), sizeof(double) * rr.size(), hipMemcpyHostToDevice));
    }
    m->max_nnz = std::max(m->max_nnz, (int)part.size());
    HIPCHK(hipMemcpy(m->pos + (size_t)w * N * 3, pos, sizeof(double) * n * 3, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->nbead + w, &n, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->emin + w, &emin, sizeof(double), hipMemcpyHostToDevice));
    m->hn[w] = n;
    {   // FNV-1a over the chain's identity: bead count, CSR offsets, partners, r0
        uint64_t h = 1469598103934665603ull;
        auto mix = [&h](const void *p, size_t nb) {
            const unsigned char *c = (const unsigned char *)p;
            for (size_t i = 0; i < nb; i++) { h ^= c[i]; h *= 1099511628211ull; }
        };
        mix(&n, sizeof n);
        mix(off.data(), sizeof(int) * off.size());
        if (!part.empty()) { mix(part.data(), sizeof(int) * part.size()); mix(rr.data(), sizeof(double) * rr.size()); }
        m->chain_sig[w] = h;
    }
    // current energy of the walker
    std::vector<double> e(m->p.nwalk);
    int r = ttc_mc_energy(m, e.data());
    if (r) return r;
    HIPCHK(hipMemcpy(m->ecur + w, &e[w], sizeof(double), hipMemcpyHostToDevice));
    return TTC_OK;
}

extern "C" int ttc_mc_get_walker(ttc_mc *m, int w, double *pos, double *energy) {
    MCHECK(m && w >= 0 && w < m->p.nwalk, TTC_ERR_ARG, "ttc_mc_get_walker: bad argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    if (pos) HIPCHK(hipMemcpy(pos, m->pos + (size_t)w * m->p.maxn * 3, sizeof(double) * m->hn[w] * 3, hipMemcpyDeviceToHost));
    if (energy) HIPCHK(hipMemcpy(energy, m->ecur + w, sizeof(double), hipMemcpyDeviceToHost));
    return TTC_OK;
}

extern "C" int ttc_mc_get_streams(ttc_mc *m, int w, uint64_t *state) {
    MCHECK(m && state && w >= 0 && w < m->p.nwalk, TTC_ERR_ARG, "ttc_mc_get_streams: bad argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(state, m->rng + (size_t)w * MC_STREAMS * 4, sizeof(uint64_t) * MC_STREAMS * 4, hipMemcpyDeviceToHost));
    return TTC_OK;
}

extern "C" int ttc_mc_set_streams(ttc_mc *m, int w, const uint64_t *state) {
    MCHECK(m && state && w >= 0 && w < m->p.nwalk, TTC_ERR_ARG, "ttc_mc_set_streams: bad argument");
    for (int s = 0; s < MC_STREAMS; s++) {
        const uint64_t *q = state + 4 * s;
        MCHECK(q[0] | q[1] | q[2] | q[3], TTC_ERR_ARG, "ttc_mc_set_streams: all-zero stream state");
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(m->rng + (size_t)w * MC_STREAMS * 4, state, sizeof(uint64_t) * MC_STREAMS * 4, hipMemcpyHostToDevice));
    return TTC_OK;
}

extern "C" int ttc_mc_energy(ttc_mc *m, double *e) {
    MCHECK(m && e, TTC_ERR_ARG, "ttc_mc_energy: bad argument");
    double *d;
    HIPCHK(hipMalloc(&d, sizeof(double) * m->p.nwalk));
    size_t shm = (size_t)(3 * m->p.maxn + MC_WAVES) * sizeof(double);
    k_mc_energy<<<m->p.nwalk, MC_THREADS, shm, m->stream>>>(mc_args(m), m->pos, m->nbead, m->noff, m->npart,
                                                            m->nr0, m->ncap, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(e, d, sizeof(double) * m->p.nwalk, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    hipFree(d);
    return TTC_OK;
}

extern "C" int ttc_mc_delta_energy(ttc_mc *m, int w, int lb, int ub, const double *newpos, double *de) {
    MCHECK(m && newpos && de && w >= 0 && w < m->p.nwalk && lb >= 0 && ub <= m->hn[w] && lb < ub, TTC_ERR_ARG,
           "ttc_mc_delta_energy: bad argument");
    double *d, *np;
    HIPCHK(hipMalloc(&d, sizeof(double)));
    HIPCHK(hipMalloc(&np, sizeof(double) * 3 * (ub - lb)));
    HIPCHK(hipMemcpy(np, newpos, sizeof(double) * 3 * (ub - lb), hipMemcpyHostToDevice));
    size_t shm = (size_t)(6 * m->p.maxn + MC_WAVES) * sizeof(double);
    k_mc_delta<<<1, MC_THREADS, shm, m->stream>>>(mc_args(m), w, lb, ub, m->pos, m->nbead, m->noff, m->npart,
                                                  m->nr0, m->ncap, np, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(de, d, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    hipFree(d);
    hipFree(np);
    return TTC_OK;
}

// ttc_mc_delta_energy without allocations: the same kernel, its input and
// output staged in this GPU's scratch array.  Nothing in the array needs to
// start at zero: the kernel reads only the coordinates copied in by this
// call and writes the result before it is read back.
extern "C" int ttc_mc_delta_energy_noalloc(ttc_mc *m, int w, int lb, int ub, const double *newpos, double *de) {
    MCHECK(m && newpos && de && w >= 0 && w < m->p.nwalk && lb >= 0 && ub <= m->hn[w] && lb < ub, TTC_ERR_ARG,
           "ttc_mc_delta_energy_noalloc: bad argument");
    MCHECK(ub - lb <= TTC_MC_NOALLOC_MAXBEADS, TTC_ERR_ARG,
           "ttc_mc_delta_energy_noalloc: more than TTC_MC_NOALLOC_MAXBEADS (1024) moved beads");
    MCHECK(m->dev >= 0 && m->dev < NOALLOC_MAXDEV && g_scratch_ptr[m->dev], TTC_ERR_ARG,
           "ttc_mc_delta_energy_noalloc: scratch array not set up for this device");
    double *sp = g_scratch_ptr[m->dev];
    HIPCHK(hipMemcpyAsync(sp + NOALLOC_STAGE, newpos, sizeof(double) * 3 * (ub - lb), hipMemcpyHostToDevice,
                          m->stream));
    size_t shm = (size_t)(6 * m->p.maxn + MC_WAVES) * sizeof(double);
    k_mc_delta<<<1, MC_THREADS, shm, m->stream>>>(mc_args(m), w, lb, ub, m->pos, m->nbead, m->noff, m->npart,
                                                  m->nr0, m->ncap, sp + NOALLOC_STAGE, sp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(de, sp, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return TTC_OK;
}

extern "C" int ttc_mc_run(ttc_mc *m, int ntrials, int mode, double lnf, int log) {
    MCHECK(m && ntrials >= 0 && (mode == 0 || mode == 1), TTC_ERR_ARG, "ttc_mc_run: bad argument");
    double escale = 2.5;
    // timing events destroyed on every return path
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() { if (e) hipEventDestroy(e); }
    } g0, g1;
    HIPCHK(hipEventCreate(&g0.e));
    HIPCHK(hipEventCreate(&g1.e));
    hipEvent_t e0 = g0.e, e1 = g1.e;
    HIPCHK(hipEventRecord(e0, m->stream));
    // at least half the CU's LDS per workgroup: one walker per CU, so no CU
    // holds two walkers while another holds none
    size_t shm = std::max(run_shm(m, mode == 1), (size_t)MC_MAX_LDS / 2 + 1024);
    MCHECK(run_shm(m, mode == 1) <= MC_MAX_LDS, TTC_ERR_ARG, "ttc_mc_run: walkers too large for the workgroup's LDS");
    if (mode == 1) {
        HIPCHK(hipFuncSetAttribute((const void *)k_mc_run<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        k_mc_run<true><<<m->p.nwalk, MC_THREADS, shm, m->stream>>>(
            mc_args(m), ntrials, lnf, log, m->pos, m->nbead, m->noff, m->npart, m->nr0, m->ncap, m->max_nnz, m->emin,
            m->ecur, m->lng, m->hist, m->rng, m->stats, m->mlp, m->logx, m->logde, m->logn, m->logcap, escale);
    } else {
        HIPCHK(hipFuncSetAttribute((const void *)k_mc_run<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        k_mc_run<false><<<m->p.nwalk, MC_THREADS, shm, m->stream>>>(
            mc_args(m), ntrials, lnf, log, m->pos, m->nbead, m->noff, m->npart, m->nr0, m->ncap, m->max_nnz, m->emin,
            m->ecur, m->lng, m->hist, m->rng, m->stats, m->mlp, m->logx, m->logde, m->logn, m->logcap, escale);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, m->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    m->t_last = 1000.0 * ms;
    return TTC_OK;
}

extern "C" int ttc_mc_last_time(ttc_mc *m, double *usec) {
    MCHECK(m && usec, TTC_ERR_ARG, "ttc_mc_last_time: bad argument");
    *usec = m->t_last;
    return TTC_OK;
}

extern "C" int ttc_mc_checkpoint(ttc_mc *m, int restore) {
    MCHECK(m && (restore == 0 || restore == 1), TTC_ERR_ARG, "ttc_mc_checkpoint: bad argument");
    const size_t W = m->p.nwalk, N = m->p.maxn, B = m->p.nbins;
    const size_t sz[5] = {sizeof(double) * W * N * 3, sizeof(double) * W, sizeof(double) * W * B,
                          sizeof(long long) * W * B, sizeof(uint64_t) * W * MC_STREAMS * 4};
    void *live[5] = {m->pos, m->ecur, m->lng, m->hist, m->rng};
    void **copy[5] = {(void **)&m->ck_pos, (void **)&m->ck_ecur, (void **)&m->ck_lng, (void **)&m->ck_hist,
                      (void **)&m->ck_rng};
    MCHECK(!restore || m->ck_pos, TTC_ERR_ARG, "ttc_mc_checkpoint: no checkpoint taken");
    for (int i = 0; i < 5; i++) {
        if (!*copy[i]) HIPCHK(hipMalloc(copy[i], sz[i]));
        if (restore) HIPCHK(hipMemcpyAsync(live[i], *copy[i], sz[i], hipMemcpyDeviceToDevice, m->stream));
        else HIPCHK(hipMemcpyAsync(*copy[i], live[i], sz[i], hipMemcpyDeviceToDevice, m->stream));
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    return TTC_OK;
}

extern "C" int ttc_mc_get_stats(ttc_mc *m, int64_t *stats) {
    MCHECK(m && stats, TTC_ERR_ARG, "ttc_mc_get_stats: bad argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(stats, m->stats, sizeof(long long) * m->p.nwalk * TTC_MC_NSTAT, hipMemcpyDeviceToHost));
    return TTC_OK;
}

extern "C" int ttc_mc_reset_stats(ttc_mc *m) {
    MCHECK(m, TTC_ERR_ARG, "ttc_mc_reset_stats: null");
    HIPCHK(hipMemsetAsync(m->stats, 0, sizeof(long long) * m->p.nwalk * TTC_MC_NSTAT, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return TTC_OK;
}

// swap_mpi.c:24-188 as one kernel over the batch: block p takes the pair
// (w, w+1), w = 2p + parity.  The sender's request (energy, window) and the
// receiver's answer (accept, energy) of msg_swap_request / msg_swap_accept
// (defines.h:1846-1874) are reads of the partner's device state.
__global__ void k_mc_swap(int parity, int nwalk, int maxn, int nbins, double ebin, double lnf, double *pos,
                          const int *nbead, const double *emin, double *ecur, double *lng, long long *hist,
                          int *accepted) {
    const int w = 2 * blockIdx.x + parity;
    if (w + 1 >= nwalk) return;
    __shared__ int acc;
    const double e1 = ecur[w], e2 = ecur[w + 1];
    const double top = nbins * ebin;
    if (threadIdx.x == 0) {
        // swap_mpi.c:147-149: each energy inside the other's window
        acc = (e2 >= emin[w] && e2 < emin[w] + top) && (e1 >= emin[w + 1] && e1 < emin[w + 1] + top);
    }
    __syncthreads();
    if (acc) {
        double *p1 = pos + (size_t)w * maxn * 3, *p2 = pos + (size_t)(w + 1) * maxn * 3;
        for (int i = threadIdx.x; i < 3 * nbead[w]; i += blockDim.x) {
            double t = p1[i];
            p1[i] = p2[i];
            p2[i] = t;
        }
    }
    if (threadIdx.x == 0) {
        const double ew[2] = {acc ? e2 : e1, acc ? e1 : e2};
        for (int s = 0; s < 2; s++) {
            // the trial counts in the bin of the walker's energy after the
            // attempt (swap_mpi.c:78-80, 131-133, 175-177), binned as the
            // walkers' own trials are
            const int v = w + s;
            const int b = (int)floor((ew[s] - emin[v]) * (1.0 / ebin));
            if (b >= 0 && b < nbins) {
                hist[(size_t)v * nbins + b] += 1;
                lng[(size_t)v * nbins + b] += lnf;
            }
            ecur[v] = ew[s];
            if (accepted) accepted[v] = acc;
        }
    }
}

extern "C" int ttc_mc_swap_windows(ttc_mc *m, int parity, double lnf, int *accepted) {
    MCHECK(m && (parity == 0 || parity == 1) && lnf >= 0.0, TTC_ERR_ARG, "ttc_mc_swap_windows: bad argument");
    const int W = m->p.nwalk;
    for (int w = parity; w + 1 < W; w += 2)
        MCHECK(m->hn[w] > 0 && m->chain_sig[w] == m->chain_sig[w + 1], TTC_ERR_ARG,
               "ttc_mc_swap_windows: neighbouring windows must hold the same chain");
    const int npair = (W - parity) / 2;
    if (npair <= 0) return TTC_OK;
    int *dacc = nullptr;
    if (accepted) HIPCHK(hipMallocAsync((void **)&dacc, sizeof(int) * W, m->stream));
    if (accepted) HIPCHK(hipMemsetAsync(dacc, 0, sizeof(int) * W, m->stream));
    k_mc_swap<<<npair, 256, 0, m->stream>>>(parity, W, m->p.maxn, m->p.nbins, m->p.ebin, lnf, m->pos, m->nbead,
                                             m->emin, m->ecur, m->lng, m->hist, dacc);
    HIPCHK(hipGetLastError());
    if (accepted) {
        HIPCHK(hipMemcpyAsync(accepted, dacc, sizeof(int) * W, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipFreeAsync(dacc, m->stream));
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    return TTC_OK;
}

// swap_mpi.c:190-296 merge_gofe(k, e_begin, e_end, mod_f): walker w is the
// left window (the sender's dos_hist, received as dos_hist_temp), w + 1 the
// right one.  ln g of both is made to agree on the overlap: the slopes of
// the two windows' ln g (central differences, one-sided at the overlap's
// ends) are averaged and integrated from the bin before the overlap in the
// left window; the right window's bins above the overlap are shifted to
// join; each adjusted bin's histogram gains (int)(change / lnf), floored at
// zero.  The windows must share the bin grid, overlap by 2 to nbins - 1
// bins and start the overlap at least one bin into the left window (the
// reference reads past the histogram otherwise).
extern "C" int ttc_mc_merge_windows(ttc_mc *m, int w, double lnf) {
    MCHECK(m && w >= 0 && w + 1 < m->p.nwalk && lnf > 0.0, TTC_ERR_ARG, "ttc_mc_merge_windows: bad argument");
    const int B = m->p.nbins;
    const double eb = m->p.ebin;
    double em[2];
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(em, m->emin + w, sizeof(double) * 2, hipMemcpyDeviceToHost));
    const double e_begin = em[0], e_end = em[0] + B * eb;
    const int e_bins = (int)((e_end - e_begin) / eb + 0.5);
    const int o_bins = (int)((e_end - em[1]) / eb);
    const int j = e_bins - o_bins;
    MCHECK(B >= 3 && o_bins >= 2 && o_bins < B && j >= 1, TTC_ERR_ARG,
           "ttc_mc_merge_windows: windows must overlap by 2 to nbins - 1 bins, the right one starting inside the left");
    std::vector<double> gl(B), gr(B);
    std::vector<long long> hl(B), hr(B);
    HIPCHK(hipMemcpy(gl.data(), m->lng + (size_t)w * B, sizeof(double) * B, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gr.data(), m->lng + (size_t)(w + 1) * B, sizeof(double) * B, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hl.data(), m->hist + (size_t)w * B, sizeof(long long) * B, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hr.data(), m->hist + (size_t)(w + 1) * B, sizeof(long long) * B, hipMemcpyDeviceToHost));
    std::vector<double> tl(o_bins), tr(o_bins), tm(o_bins);
    for (int i = 0; i < o_bins; i++) {
        // slopes over two bins; the left window's last bin repeats the slope
        // before it, the right window's first bin looks two bins ahead
        tl[i] = (i == o_bins - 1) ? tl[i - 1] : gl[j + i + 1] - gl[j + i - 1];
        tr[i] = (i == 0) ? gr[2] - gr[0] : gr[i + 1] - gr[i - 1];
        tm[i] = (tl[i] + tr[i]) / 4.0;
    }
    for (int i = 0; i < o_bins; i++) {
        const double old = gl[j + i];
        gl[j + i] = gl[j + i - 1] + tm[i];
        hl[j + i] = std::max(0LL, hl[j + i] + (long long)(int)((gl[j + i] - old) / lnf));
    }
    const double shift = gr[o_bins - 1] - gl[j + o_bins - 1];
    for (int i = o_bins; i < B; i++) gr[i] -= shift;
    for (int i = 0; i < o_bins; i++) {
        const double old = gr[i];
        gr[i] = gl[j + i];
        hr[i] = std::max(0LL, hr[i] + (long long)(int)((gr[i] - old + shift) / lnf));
    }
    HIPCHK(hipMemcpy(m->lng + (size_t)w * B, gl.data(), sizeof(double) * B, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->lng + (size_t)(w + 1) * B, gr.data(), sizeof(double) * B, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->hist + (size_t)w * B, hl.data(), sizeof(long long) * B, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->hist + (size_t)(w + 1) * B, hr.data(), sizeof(long long) * B, hipMemcpyHostToDevice));
    return TTC_OK;
}

extern "C" int ttc_mc_get_dos(ttc_mc *m, int w, double *lng, int64_t *hist) {
    MCHECK(m && w >= 0 && w < m->p.nwalk, TTC_ERR_ARG, "ttc_mc_get_dos: bad argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    if (lng) HIPCHK(hipMemcpy(lng, m->lng + (size_t)w * m->p.nbins, sizeof(double) * m->p.nbins, hipMemcpyDeviceToHost));
    if (hist) HIPCHK(hipMemcpy(hist, m->hist + (size_t)w * m->p.nbins, sizeof(long long) * m->p.nbins, hipMemcpyDeviceToHost));
    return TTC_OK;
}

extern "C" int ttc_mc_set_dos(ttc_mc *m, int w, const double *lng) {
    MCHECK(m && lng && w >= 0 && w < m->p.nwalk, TTC_ERR_ARG, "ttc_mc_set_dos: bad argument");
    HIPCHK(hipMemcpy(m->lng + (size_t)w * m->p.nbins, lng, sizeof(double) * m->p.nbins, hipMemcpyHostToDevice));
    return TTC_OK;
}

extern "C" int ttc_mc_reset_hist(ttc_mc *m) {
    MCHECK(m, TTC_ERR_ARG, "ttc_mc_reset_hist: null");
    HIPCHK(hipMemsetAsync(m->hist, 0, sizeof(long long) * m->p.nwalk * m->p.nbins, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return TTC_OK;
}

extern "C" int ttc_mc_get_log(ttc_mc *m, int max, float *features, double *de, int *count) {
    MCHECK(m && count, TTC_ERR_ARG, "ttc_mc_get_log: bad argument");
    int n = 0;
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(&n, m->logn, sizeof(int), hipMemcpyDeviceToHost));
    n = std::min(n, m->logcap);
    *count = n;
    int k = std::min(n, max);
    if (k > 0 && features) HIPCHK(hipMemcpy(features, m->logx, sizeof(float) * (size_t)k * NF, hipMemcpyDeviceToHost));
    if (k > 0 && de) HIPCHK(hipMemcpy(de, m->logde, sizeof(double) * k, hipMemcpyDeviceToHost));
    return TTC_OK;
}

extern "C" int ttc_mc_clear_log(ttc_mc *m) {
    MCHECK(m, TTC_ERR_ARG, "ttc_mc_clear_log: null");
    HIPCHK(hipMemsetAsync(m->logn, 0, sizeof(int), m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return TTC_OK;
}

extern "C" int ttc_mc_set_mlp(ttc_mc *m, const float *params) {
    MCHECK(m && params, TTC_ERR_ARG, "ttc_mc_set_mlp: bad argument");
    HIPCHK(hipMemcpy(m->mlp, params, sizeof(float) * NPAR, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m->adam, 0, sizeof(float) * 2 * NPAR));
    m->adam_t = 0;
    return TTC_OK;
}

extern "C" int ttc_mc_get_mlp(ttc_mc *m, float *params) {
    MCHECK(m && params, TTC_ERR_ARG, "ttc_mc_get_mlp: bad argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(params, m->mlp, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
    return TTC_OK;
}

extern "C" int ttc_mlp_forward(ttc_mc *m, int nsamp, const float *x, float *out) {
    MCHECK(m && x && out && nsamp > 0, TTC_ERR_ARG, "ttc_mlp_forward: bad argument");
    float *dx, *dy;
    HIPCHK(hipMalloc(&dx, sizeof(float) * (size_t)nsamp * NF));
    HIPCHK(hipMalloc(&dy, sizeof(float) * nsamp));
    HIPCHK(hipMemcpy(dx, x, sizeof(float) * (size_t)nsamp * NF, hipMemcpyHostToDevice));
    k_mlp_forward<<<1, TR_THREADS, train_shm(), m->stream>>>(m->mlp, nsamp, dx, dy);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dy, sizeof(float) * nsamp, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    hipFree(dx);
    hipFree(dy);
    return TTC_OK;
}

extern "C" int ttc_mlp_loss_grad(ttc_mc *m, int nsamp, const float *x, const float *y, double *loss, float *grad) {
    MCHECK(m && x && y && loss && grad && nsamp > 0, TTC_ERR_ARG, "ttc_mlp_loss_grad: bad argument");
    float *dx, *dy, *dg;
    double *dl;
    HIPCHK(hipMalloc(&dx, sizeof(float) * (size_t)nsamp * NF));
    HIPCHK(hipMalloc(&dy, sizeof(float) * nsamp));
    HIPCHK(hipMalloc(&dg, sizeof(float) * NPAR));
    HIPCHK(hipMalloc(&dl, sizeof(double)));
    HIPCHK(hipMemcpy(dx, x, sizeof(float) * (size_t)nsamp * NF, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dy, y, sizeof(float) * nsamp, hipMemcpyHostToDevice));
    k_mlp_loss_grad<<<1, TR_THREADS, train_shm(), m->stream>>>(m->mlp, nsamp, dx, dy, dl, dg);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, dg, sizeof(float) * NPAR, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(loss, dl, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    hipFree(dx); hipFree(dy); hipFree(dg); hipFree(dl);
    return TTC_OK;
}

extern "C" int ttc_mlp_train(ttc_mc *m, int nepoch, int batch, float lr, double escale, double *loss) {
    MCHECK(m && loss && nepoch >= 0, TTC_ERR_ARG, "ttc_mlp_train: bad argument");
    (void)batch;   // the minibatch is fixed at 32 moves (64 network rows)
    int n = 0;
    HIPCHK(hipMemcpy(&n, m->logn, sizeof(int), hipMemcpyDeviceToHost));
    n = std::min(n, m->logcap);
    MCHECK(n >= TR_B, TTC_ERR_ARG, "ttc_mlp_train: fewer logged moves than one minibatch");
    double *dl;
    HIPCHK(hipMalloc(&dl, sizeof(double)));
    k_mlp_train<<<1, TR_THREADS, train_shm(), m->stream>>>(m->mlp, m->adam, m->adam_t, n, m->logx, m->logde, nepoch,
                                                           lr, escale, 2.0 * m->p.nbins * m->p.ebin,
                                                           m->p.seed + (uint64_t)m->adam_t, dl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(loss, dl, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->adam_t += nepoch * (n / TR_B);
    hipFree(dl);
    return TTC_OK;
}
