// ttc_common.h -- device state and shared helpers of libttc.so (gfx950).
//
// Data layout in HBM (one context = one periodic box):
//   Per-site state is kept in "sorted order": sites ordered by neighbour
//   cell, and by original index inside a cell.  The order is rebuilt with
//   every neighbour-list build and all per-site arrays are permuted with it,
//   so every per-step kernel reads and writes coalesced.
//     pos4[s]  = {x, y, z, q}            FP64 (positions wrapped to [-L/2,L/2])
//     vel4[s]  = {vx, vy, vz, 100/m}     FP64 (A/ps; 100/m turns kJ/mol/A into A/ps^2)
//     f4[s]    = {fx, fy, fz, 0}         FP64 total force of the last evaluation
//     type[s], orig[s] (original index), inv[orig] = s
//     anchor[s] = position at the last list build (DLIST)
//     fixj[3][N] int64 partner forces of the half-list pair kernel (2^-24 kJ/mol/A)
//   Neighbour list: full list (each pair in both rows) or, in
//   velocity-Verlet loops, half list (each pair in the row of its lower site,
//   entries holding window slots instead of j; see ttc_ctx::half_mode), ELL
//   layout per group of 64 consecutive sorted sites, tiled by 16 entries:
//   nbl[g][k/16][lane][k%16] (uint32), glen[g] entries per row (cap, the
//   row capacity, is a multiple of 64).  A row's consecutive entries share a
//   64-byte segment, so appends by the list build coalesce, and the 16 k of
//   one 4 KB tile are what the pair kernel's waves read together.
//   Entry = j (sorted index, 24 bits) | type of j << 24 (6 bits) |
//   class << 30 (0 plain, 1 excluded xflag 1/2/3, 2 one-four xflag 4);
//   0xFFFFFFFF pads short rows.  Carrying j's type in the entry spares the
//   pair kernel a second gather per entry.
//   PME: charge grid as int64 fixed point (2^-40 e units) for deterministic
//   atomic spreading, complex FP32 work grid, FP32 influence function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ttc.h"

#define TTC_WAVE 64
#define TTC_PAD 0xFFFFFFFFu
#define TTC_JMASK 0x00FFFFFFu
#define TTC_TSHIFT 24
#define TTC_MAX_SITES (1 << 24)
#define TTC_FIX_SCALE 4294967296.0          /* 2^32 */
#define TTC_FIX_INV (1.0 / 4294967296.0)
#define TTC_FIXJ_SHIFT 8                    /* partner forces in 2^-24 units: << 8 gives 2^-32 */
#define TTC_QFIX_SCALE 1099511627776.0      /* 2^40 */
#define TTC_QFIX_INV (1.0 / 1099511627776.0)
#define TTC_MAX_TYPES 64
#define TTC_VV_RUN 10                       /* plain MD steps captured in one graph */

// defines.h:60-71, product in the same order as the reference macro
#define TTC_PI 3.141592653589793238462643383
#define TTC_KB 1.380662E-23
#define TTC_NA 6.022045E23
#define TTC_AU 1.6605655E-27
#define TTC_QFACTOR (8.987551789e19 * 1.602176462e-19 * 1.602176462e-19 * 6.022045E23 * 1.0E-3)

enum { ST_NLFLAG = 0, ST_OVERFLOW = 1, ST_MAXROW = 2, ST_GRIDERR = 3, ST_NSTAT = 12 };
// ST_OVERFLOW bits: 1 a row longer than the list capacity, 2 a cell fuller
// than its capacity, 4 an exported row longer than max_nb, 8 a half-list
// window wider than its capacity (or a partner beyond the builder's bitmap),
// 16 a pair force beyond the range of the partner accumulators.  Bits 1 and
// 8 stop an MD loop's remaining steps until the host has grown the list.
#define TTC_OVF_REDO (1 | 8)

// energy/virial slots reduced on the device (FP64)
enum {
    EV_LJ = 0, EV_LJS, EV_REAL, EV_REALS, EV_INTRA,   // pair loop
    EV_PV0, EV_PV1, EV_PV2, EV_PV3, EV_PV4, EV_PV5,   // pair virial
    EV_REC, EV_KV0, EV_KV1, EV_KV2, EV_KV3, EV_KV4, EV_KV5,  // reciprocal
    EV_BOND, EV_BV0, EV_BV1, EV_BV2, EV_BV3, EV_BV4, EV_BV5, // bonds
    EV_KIN, EV_KE0, EV_KE1, EV_KE2, EV_KE3, EV_KE4, EV_KE5,  // kinetic tensor
    EV_FSX, EV_FSY, EV_FSZ,                           // sum of PME forces
    // CONFIGT: sums over sites of d2U/dx_i2, dy_i2, dz_i2 (Hessian diagonal)
    EV_PHX, EV_PHY, EV_PHZ,                           // pair terms
    EV_KHX, EV_KHY, EV_KHZ,                           // reciprocal energy
    EV_BHX, EV_BHY, EV_BHZ,                           // bonds
    // angle-type terms: energies per kind, virial and Hessian diagonal of
    // the kinds a launch evaluated
    EV_EBEND, EV_EUREY, EV_ETORS, EV_EIMPR,
    EV_AV0, EV_AV1, EV_AV2, EV_AV3, EV_AV4, EV_AV5,
    EV_AHX, EV_AHY, EV_AHZ,
    EV_MOX, EV_MOY, EV_MOZ,                           // net momentum (drift.c)
    EV_N
};

// bonded-term selection of a gather launch: bonds and the angle-type kinds
// (cbend.c, curey.c, ctorsion.c, cimproper.c); force_short.c evaluates the
// first three, force_long.c the last two
#define TTC_BT_BOND 1
#define TTC_BT_BEND 2
#define TTC_BT_UREY 4
#define TTC_BT_TORS 8
#define TTC_BT_IMPR 16
#define TTC_BT_ALL 31
#define TTC_BT_SHORT (TTC_BT_BOND | TTC_BT_BEND | TTC_BT_UREY)
#define TTC_BT_LONG (TTC_BT_TORS | TTC_BT_IMPR)

struct FftPlan {
    int n = 0;
    int nrad = 0;
    int rad[16];
    float2 *tw = nullptr;   // e^{+2 pi i m / n}, m = 0..n-1
};

// x 2^32 / L per axis, wrapped modulo 2^32 (modulo the box), and the
// charge's FP32 bits: the half-list pair kernel's coordinates; the 32-bit
// difference of two of them is the minimum-image displacement
__device__ __forceinline__ int4 ttc_posq(const double4 &p, double kx, double ky, double kz) {
    return make_int4((int)(unsigned)__double2ll_rn(p.x * kx), (int)(unsigned)__double2ll_rn(p.y * ky),
                     (int)(unsigned)__double2ll_rn(p.z * kz), __float_as_int((float)p.w));
}

struct ttc_ctx {
    ttc_params p{};
    int n = 0, G = 0, cap = 0, ntypes = 0;
    int order = 4;
    double e_self = 0.0;
    bool topo = false;
    bool have_list = false;
    bool list_pruned = false;   // list built in an MD loop at the pruned cutoff
    hipStream_t stream = nullptr;
    // second stream of the captured MD steps: the FFTs and the gather run on
    // it, concurrently with the list rebuild and the pair kernel (fork after
    // the drift, join before the kick); sev[0..2] mark the fork points
    hipStream_t side = nullptr;
    hipEvent_t sev[3] = {nullptr, nullptr, nullptr};

    // sorted-order state
    double4 *pos4 = nullptr, *vel4 = nullptr, *f4 = nullptr, *anchor = nullptr;
    double4 *pos4b = nullptr, *vel4b = nullptr, *f4b = nullptr, *anchorb = nullptr;
    float4 *posfb = nullptr;        // FP32 positions in the new order (fast list build)
    int *type = nullptr, *typeb = nullptr, *orig = nullptr, *origb = nullptr, *inv = nullptr;
    long long *fix = nullptr;

    // cells
    int ncx = 0, ncy = 0, ncz = 0, ncell = 0;
    int *cell_of = nullptr, *cell_cnt = nullptr, *cell_start = nullptr, *cell_acc = nullptr;
    int *cell_members = nullptr, *perm = nullptr;

    // list
    uint32_t *nbl = nullptr;
    int *glen = nullptr;
    int *gcnt = nullptr;            // entries (non-pad) per group of the last build
    int *status = nullptr;          // device ST_* flags
    int *h_status = nullptr;        // pinned host mirror

    // exclusions, original numbering, CSR both directions
    int *ex_off = nullptr, *ex_part = nullptr, *ex_flag = nullptr;
    int ex_span = 0;                // max |a-b| over exclusions
    int nx = 0;

    // LJ tables: {eps, sig^2, eps14, sig14^2} per type pair
    float4 *ljtab = nullptr;

    // bonds, per site in original numbering (CSR, both directions)
    int nbond = 0;
    int *b_off = nullptr, *b_part = nullptr;
    double *b_r0 = nullptr, *b_k = nullptr;

    // force contributions of one evaluation (sorted order)
    double4 *fpair = nullptr;       // pair kernel
    double4 *g4 = nullptr;          // PME gather + bonds
    // MD step: positions + charge by original index (written by the drift)
    // and the gather's output by original index, so the gather does not
    // wait for the step's rebuild
    double4 *snap = nullptr, *g4o = nullptr;
    // MD step: the spread's inputs as the drift leaves them (positions +
    // charge in the sorted order, the cell ranges of that order), read by
    // the spread on the side stream while a rebuild rewrites the originals
    double4 *spos = nullptr;
    int *scs = nullptr, *scc = nullptr;
    bool f4_valid = false;          // f4 holds the force at the current positions
    // RESPA short (bonded) and long (nonbonded) forces, sorted order, permuted
    // with every rebuild like f4
    double4 *fs = nullptr, *fl = nullptr, *fsb = nullptr, *flb = nullptr;
    bool mts_valid = false;

    // Nose-Hoover chain (integrate_nhc.c): device state {r,v,G}*M (M <= 8),
    // then the velocity scale at [TTC_NHC_SCALE], target T and mass Q at
    // [TTC_NHC_T], [TTC_NHC_Q] (device-resident so captured graphs follow
    // temperature changes), and a scratch scale factor at [TTC_NHC_VMUL]
    double *nhc = nullptr;
    double nhc_T = 300.0, nhc_Q = 1.0, nhc_nfree = 0.0, nhc_w[8] = {1, 0, 0, 0, 0, 0, 0, 0};
    int nhc_M = 0, nhc_Nc = 1, nhc_Nys = 1;
    hipGraphExec_t g_nhc_plain = nullptr, g_nhc_energy = nullptr;
    hipGraphExec_t g_mts_plain = nullptr, g_mts_energy = nullptr;
    int g_mts_ninner = 0;
    double g_mts_dtlong = 0.0, g_mts_vscal = 0.0;

    // PME
    int K[3] = {0, 0, 0};
    size_t ngrid = 0;
    long long *qfix = nullptr;
    float2 *qc = nullptr;
    float *qr = nullptr;
    float *eterm = nullptr;
    FftPlan plan[3];
    // theta(d) for grid offsets |d_x|,|d_y|,|d_z| < order: the real-space
    // kernel of the reciprocal energy near the origin (CONFIGT Hessian)
    float *theta_near = nullptr;

    // angle-type bonded terms, one thread per term, kinds in this order
    // (bend, urey, tors, impr), each kind's range padded to whole waves:
    // ang_start[k] is kind k's first term (ang_start[4] the padded total),
    // ang_slot[k] its first force slot (3, 2, 4, 4 slots per term: one per
    // site); ang_tf[slot] holds a term's force on one of its sites, which
    // the gather adds per site through the CSR ang_off / ang_ref
    // (slot | kind << 30) in the original site order
    int nbend = 0, nurey = 0, ntors = 0, nimpr = 0;
    int ang_start[5] = {0, 0, 0, 0, 0};
    int ang_slot[5] = {0, 0, 0, 0, 0};
    int4 *ang_idx = nullptr;        // [ang_start[4]] sites (x < 0: padding)
    double2 *ang_par = nullptr;     // [ang_start[4]] equilibrium, force constant
    double *ang_tpar = nullptr;     // [ntors][18] kphi[6], nphi[6], delphi[6]
    double4 *ang_tf = nullptr;      // [ang_slot[4]]
    int *ang_off = nullptr;         // [n + 1]
    int *ang_ref = nullptr;

    // drift.c: remove the net momentum after every drift_every-th MD step
    // (nve_md.c:208, nvt_md.c:302); 0: never
    int drift_every = 0;

    // reductions: partials[slot * nblk_max + block]
    double *partials = nullptr;
    int nblk_max = 0;
    double *ev = nullptr;           // [EV_N] reduced
    double *h_ev = nullptr;         // pinned
    double *rows = nullptr;         // energy rows of ttc_md_vverlet
    int rows_cap = 0;

    // graph; gen counts reallocations of buffers that captured graphs
    // reference (list, energy rows), so holders of graphs can re-capture
    // velocity-Verlet step graphs [head][energy]: head = the previous
    // step's second half kick fused into this step's drift
    hipGraphExec_t g_vv[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    hipGraphExec_t g_vv_run = nullptr;   // TTC_VV_RUN plain steps with the deferred kick, one graph
    unsigned gen = 0;
    double g_vscal = 1.0;

    // tuned launch shape: waves per 64-site group of the pair kernel (4, 8, 16)
    int pair_waves = 8;

    // Half list (velocity-Verlet MD loops): each pair once, in the row of
    // the lower site (minimum-image dz > 0, ties by index).  An entry holds the slot of j in its group's window (the
    // group's distinct partners, ascending, win[g * wcap + slot] = j) instead
    // of j; the pair kernel adds the partners' forces into LDS int64
    // accumulators per window slot and flushes them to fixj.
    bool half_mode = false;         // conditional rebuilds build half lists
    bool list_half = false;         // the current list is a half list
    int4 *posq = nullptr;           // [N] pos4 as 32-bit fixed point + charge bits (half-list kernel),
                                    // written by every MD drift and by the MD-loop rebuild's copy
    int *win = nullptr;             // [G][wcap] window of every group
    int *wlen = nullptr;            // [G] window length
    int wcap = 0;                   // window capacity (multiple of 64)
    int span = 0;                   // partner bitmap of the builder, bits (multiple of 64)
    int span_pre = 0;               // its bit of the group's first site
    long long *fixj = nullptr;      // [3][n] partner forces, 2^-24 kJ/mol/A; read and cleared by the combine
    int pair_waves_h = 8;           // waves per group of the half-list pair kernel
    int spread_chunk = 0;           // sites per spread pass (0: the occupancy rule)
    // limits below the capacities (0: none) at which MD-loop builds report
    // an overflow; cleared by the next growth (ttc_set_capacity_limits)
    int row_limit = 0, win_limit = 0;

    // timing (microseconds)
    double t_pair = 0, t_spread = 0, t_step = 0;
};

enum { ST_NREBUILD = 4, ST_STEP = 5, ST_NROWS = 6, ST_TICKET = 7, ST_WMAX = 8, ST_SPANMAX = 9, ST_PREMAX = 10 };
enum { TTC_NHC_SCALE = 24, TTC_NHC_T = 25, TTC_NHC_Q = 26, TTC_NHC_VMUL = 27, TTC_NHC_LEN = 28 };

#define ROW_LEN 9

// ---- launchers (defined in the .hip files) ----
// nblist
int ttc_launch_rebuild(ttc_ctx *c, bool conditional);   // cells, sort, permute, list
void ttc_launch_nlcheck(ttc_ctx *c);
// unconditional group-cooperative build at the current positions (cells,
// sort, pruned list in the context's half_mode format), as an MD loop's
// rebuild makes it; the host-side redo after an overflow inside a loop
int ttc_launch_rebuild_md(ttc_ctx *c);
// a pruned list (half or full format) at the anchors in the current sorted
// order, leaving the site state alone
int ttc_launch_build_md_at_anchors(ttc_ctx *c, bool half);
// window capacity and builder bitmap size for this context's geometry
void ttc_window_sizes(const ttc_ctx *c, int *wcap, int *span, int *pre);
void ttc_bitmap_sizes(const ttc_ctx *c, int below, int above, int *span, int *pre);
int ttc_export_list(ttc_ctx *c, double cutoff, int max_nb, int *count, int *list, int *xflag);
int ttc_nl_check_range(ttc_ctx *c, int lb, int ub, int *flag);
double ttc_list_cutoff(const ttc_ctx *c);
// cutoff of lists built inside MD loops: rc + 2 (rclist-rc)/4, complete for
// every pair within rc while the DLIST trigger has not fired (each site
// within (rclist-rc)/4 of its anchor)
double ttc_compute_cutoff(const ttc_ctx *c);
int ttc_rebuild_full_at_anchors(ttc_ctx *c);   // list at the full cutoff from the anchors
int ttc_cellcap(const ttc_ctx *c);
// pair; the range variant weights each pair by whether a site lies in
// [lb,ub) of the original numbering (energies only)
void ttc_launch_pair(ttc_ctx *c, bool energy, int blk_off);
// pair kernel of an MD step: also counts the step and clears the trigger
// flag the step's conditional rebuild has consumed
void ttc_launch_pair_step(ttc_ctx *c, bool energy);
void ttc_launch_pair_drifted(ttc_ctx *c, bool energy);
// MD loops keep half lists (each pair once) from 256 sites on
inline bool ttc_half_enabled(const ttc_ctx *c) { return c->n >= 256; }
void ttc_launch_pair_range(ttc_ctx *c, int lb, int ub);
void ttc_launch_pair_hes(ttc_ctx *c);    // energies + Hessian diagonal (CONFIGT)
int ttc_pair_blocks(ttc_ctx *c);
// pme (+ bonds in the gather)
int ttc_pme_setup(ttc_ctx *c);
int ttc_pme_attrs(ttc_ctx *c);   // LDS limits of the spread kernel (needs the cell geometry)
void ttc_pme_free(ttc_ctx *c);
// gather: pme -> PME forces, bonds -> bond forces (either may be off)
void ttc_launch_pme(ttc_ctx *c, bool energy, bool pme, int bterms);   // bterms: TTC_BT_* mask
// the two halves of ttc_launch_pme: the FFT convolution of the charge grid
// (charge grid -> qr) and the gather (qr + positions -> g4)
void ttc_launch_fft(ttc_ctx *c, bool energy);
void ttc_launch_gather(ttc_ctx *c, bool energy, bool pme, int bterms, bool hes);
// angle-type terms of the kinds in bterms (TTC_BT_BEND..IMPR) at pos4 (sorted
// order, inv map) or, with snap, at the drift's snapshot (original order);
// the gather that follows on the same stream adds their forces
void ttc_launch_angles(ttc_ctx *c, bool energy, bool hes, int bterms, bool snap);
int ttc_angle_blocks(const ttc_ctx *c);
// captured MD steps: side stream waits for the work queued on the main
// stream so far (ttc_side_fork, event k), main waits for the side stream
// (ttc_side_join)
void ttc_side_fork(ttc_ctx *c, int k);
void ttc_side_join(ttc_ctx *c);
// run launches on the side stream for the lifetime of the object
struct TtcOnSide {
    ttc_ctx *c;
    hipStream_t saved;
    explicit TtcOnSide(ttc_ctx *cc) : c(cc), saved(cc->stream) { cc->stream = cc->side; }
    ~TtcOnSide() { c->stream = saved; }
};
// energies + forces + Hessian diagonal (CONFIGT); needs ttc_pme_theta first
void ttc_launch_pme_hes(ttc_ctx *c, bool pme, int bterms);
int ttc_pme_theta(ttc_ctx *c);   // fills theta_near (once per context)
int ttc_pme_conv_blocks(ttc_ctx *c);
int ttc_gather_blocks(ttc_ctx *c);
void ttc_launch_spread(ttc_ctx *c);
// drift of an MD step: writes the original-order snapshot for the gather
// and the spread's inputs; kick = the previous step's second half kick fused
// ahead of it
void ttc_launch_md_drift(ttc_ctx *c, double vscal, bool kick);
// drift of an NHC step (no fused kick) with the spread's copies
void ttc_launch_nhc_drift(ttc_ctx *c);
// spread of an MD step from the drift's copies (spos, scs, scc)
void ttc_launch_md_spread(ttc_ctx *c);
// the gather in the original numbering (snapshot in, g4o out)
void ttc_launch_gather_orig(ttc_ctx *c, bool energy);
// md / combine / reductions
// mask bits: 1 pair, 2 gather (PME + bonds), 4 remove mean PME force
void ttc_launch_combine(ttc_ctx *c, int mask);
void ttc_launch_combine_to(ttc_ctx *c, int mask, double4 *out);
void ttc_launch_kick(ttc_ctx *c, double vscal, bool energy);
// the same kick leaving the step counter and the trigger flag alone (the
// MD loop's pair kernel keeps them), on the original-order gather output
void ttc_launch_kick_tail(ttc_ctx *c, double vscal, bool energy);
// v = A v + B (100/m) F; optional trigger check, KE partials, flag clear
void ttc_launch_vkick(ttc_ctx *c, const double4 *F, double A, double B, bool check, bool ke, bool clear_flag);
// v = A v + B (100/m) F, then x += dt v and pbc (no spreading)
void ttc_launch_vkick_drift(ttc_ctx *c, const double4 *F, double A, double B, double dt);
void ttc_launch_nhc(ttc_ctx *c);         // chain update from the KE partials -> scale
void ttc_launch_vscale(ttc_ctx *c);      // v *= scale
void ttc_launch_vmul(ttc_ctx *c);        // v *= nhc[TTC_NHC_VMUL]
void ttc_launch_kinet(ttc_ctx *c);
// drift.c:20-72: v_i -= (sum_j m_j v_j / N) / m_i (two launches: block sums
// of the momentum, then every block sums them in a fixed order and corrects)
void ttc_launch_momentum_removal(ttc_ctx *c);
void ttc_launch_reduce(ttc_ctx *c, bool append_row);
int ttc_kick_blocks(ttc_ctx *c);
double ttc_nr_erfc_host(double x);

void ttc_set_error(const std::string &msg);

#define TTC_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ttc_set_error(std::string(#call) + ": " + hipGetErrorString(e_));      \
            return TTC_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

// ---- device helpers ----
__device__ __forceinline__ long long ttc_to_fix(double v) {
    return __double2ll_rn(v * TTC_FIX_SCALE);
}

// Minimum image exactly as cnonbond_nblist.c:166-173 (one-sided wrap).
// Kept as branches: almost every pair takes neither, so they are cheaper
// than computing both shifted values and selecting.
__device__ __forceinline__ double ttc_minimg(double d, double h, double l) {
    if (d > h) d -= l;
    else if (d < -h) d += l;
    return d;
}

// FP64 sum of squares rounded op by op, (dx*dx + dy*dy) + dz*dz, as the
// reference's C computes it (x86-64 without FMA).  Contraction into FMAs is
// switched off here: it would change which boundary pairs pass the test.
__device__ __forceinline__ double ttc_r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    double x2 = dx * dx;
    double y2 = dy * dy;
    double z2 = dz * dz;
    return (x2 + y2) + z2;
}

// index of entry k of row (lane) l inside a group's list block (cap*64 words)
__device__ __forceinline__ size_t ttc_nbl_idx(int k, int l) {
    return (size_t)(k >> 4) * 1024 + (size_t)l * 16 + (k & 15);
}

__device__ __forceinline__ double ttc_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// mean PME force over all sites from the gather's per-block sums, reduced
// in a fixed order (lane-strided, then a shuffle tree), wave d for
// component d; every thread of the workgroup (at least 192) must call it
// (it ends with a barrier)
__device__ __forceinline__ void ttc_mean_pme(const double *partials, int maxb, int nbg, int n, double *sh3) {
    const int d = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (d < 3) {
        double t = 0.0;
        for (int b = lane; b < nbg; b += 64) t += partials[(size_t)(EV_FSX + d) * maxb + b];
        t = ttc_wave_sum(t);
        if (lane == 0) sh3[d] = t / (double)n;
    }
    __syncthreads();
}

// The per-site force of one evaluation: pair force (mask 1) + PME and bond
// force of the gather (mask 2) - mean PME force (mask 4), added as int64
// fixed point (2^-32 kJ/mol/A) so the total does not depend on the order.
// The gather force is read at gi (the site's original index where the
// gather ran in the original numbering), the pair force at s.
// With mask 1 the partner forces of a half-list pair kernel (fixj, 2^-24
// units, [3][n]) are added too and cleared for the next evaluation (they are
// zero after a full-list pair kernel).
__device__ __forceinline__ double4 ttc_combine(int s, int mask, const double4 *fpair, const double4 *g4,
                                               const double *mean, int gi = -1, long long *fixj = nullptr,
                                               int n = 0) {
    long long fx = 0, fy = 0, fz = 0;
    if (mask & 1) {
        double4 a = fpair[s];
        fx += ttc_to_fix(a.x); fy += ttc_to_fix(a.y); fz += ttc_to_fix(a.z);
        if (fixj) {
            fx += fixj[s] << TTC_FIXJ_SHIFT;
            fy += fixj[n + s] << TTC_FIXJ_SHIFT;
            fz += fixj[2 * n + s] << TTC_FIXJ_SHIFT;
            fixj[s] = 0;
            fixj[n + s] = 0;
            fixj[2 * n + s] = 0;
        }
    }
    if (mask & 2) {
        double4 b = g4[gi < 0 ? s : gi];
        fx += ttc_to_fix(b.x); fy += ttc_to_fix(b.y); fz += ttc_to_fix(b.z);
    }
    if (mask & 4) {
        fx -= ttc_to_fix(mean[0]); fy -= ttc_to_fix(mean[1]); fz -= ttc_to_fix(mean[2]);
    }
    return make_double4((double)fx * TTC_FIX_INV, (double)fy * TTC_FIX_INV, (double)fz * TTC_FIX_INV, 0.0);
}

// ---- batched launches ------------------------------------------------------
// One grid covering the blocks of many contexts: block b runs block
// map[b].blk of context map[b].rep's own grid with that context's argument
// pack pk[map[b].rep], so a batched launch computes exactly what launching
// every context's kernel on its own would.  Stages are built once per set of
// contexts (ttc_stage_upload) and launched by the kernel files' ttc_blaunch_*.
struct TtcBmap {
    int rep, blk;
};
struct TtcStage {
    void *pk = nullptr;        // device array of per-context argument packs
    TtcBmap *map = nullptr;    // device array, one entry per block
    int nblk = 0;
    size_t shm = 0;            // dynamic LDS per block (largest over the contexts)
    int kind = 0;              // variant of the stage's kernel (pair: 1 = half lists)
};
int ttc_stage_upload(TtcStage *st, const void *packs, size_t pack_bytes, int nrep, const std::vector<int> &blocks,
                     size_t shm);
void ttc_stage_free(TtcStage *st);
template <typename Pack>
int ttc_stage_make(TtcStage *st, const std::vector<Pack> &packs, const std::vector<int> &blocks, size_t shm) {
    return ttc_stage_upload(st, packs.data(), sizeof(Pack), (int)packs.size(), blocks, shm);
}

