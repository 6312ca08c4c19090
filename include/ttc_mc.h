/*
 * ttc_mc.h -- C ABI of the batched Monte Carlo / density-of-states path
 * (libttc.so): many C-alpha Go-model walkers of uneven length sampled at
 * once on the device, Wang-Landau acceptance, exact FP64 energy changes,
 * and an optional learned first stage (delayed acceptance).
 *
 * Reference counterparts (file:line in the reference tree):
 *   pivot and displacement moves   mcmove_pivot.c:32-283, mcmove_displ.c:33-139
 *   range energy of a moved part   forces_mc.c:90-300 (o_potens / n_potens),
 *                                  enbnd_mc.c:26-504 (pairs touching [lb,ub))
 *   Go energy terms                cgolik.c:644-948, 12-10 form 1202-1229
 *   Wang-Landau acceptance         pivot.c:236-253, ctdos_pivot.c:246-314,
 *                                  ln g interpolation dos_interpolate.c:107-124
 *   random numbers                 random.c:37-75 (ran2; here a per-walker
 *                                  counter-free generator, see ttc_mc_run)
 * The learned first stage and its training have no reference counterpart.
 * Error codes are those of ttc.h.
 */
#ifndef TTC_MC_H
#define TTC_MC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ttc_mc ttc_mc;

typedef struct {
    int    nwalk;        /* walkers                                         */
    int    maxn;         /* longest chain (beads)                           */
    int    nbins;        /* Wang-Landau bins per walker                     */
    double ebin;         /* bin width, kJ/mol                               */
    double kbond;        /* bond energy kbond (r - r0bond)^2 (cbond.c:103)  */
    double r0bond;
    double eps;          /* native 12-10 depth and non-native repulsion     */
    double sig_rep;      /* non-native repulsion eps (sig_rep/r)^12         */
    double pivot_frac;   /* probability of a pivot move (else displacement) */
    double theta_max;    /* pivot angle = u * pi * theta_max                */
    double disp_max;     /* displacement = (u - 0.5) * disp_max per axis     */
    uint64_t seed;
} ttc_mc_params;

/* Surrogate input features of a pivot move (all rounded to bfloat16):
 *   [0,16)  min((sig/r)^6, 16) for the 4 moved x 4 fixed beads nearest the
 *           pivot, old positions;  [16,32) the same, new positions
 *   [32,40) the same term for the closest fixed bead (|i-j| >= 3) of the
 *           moved beads at chain offsets 2, 4, ..., 256 from the pivot, old;
 *           [40,48) new
 *   [48] 1, [49] |angle|/pi, [50] moved fraction, [51] 0
 * The reverse move swaps old and new.                                       */
#define TTC_MC_NFEAT 52  /* surrogate input features per move */
#define TTC_MC_HID   64  /* surrogate hidden width            */

/* stats per walker (ttc_mc_get_stats), int64 each */
#define TTC_MC_STAT_TRIALS   0
#define TTC_MC_STAT_ACCEPT   1
#define TTC_MC_STAT_STAGE1_REJECT 2
#define TTC_MC_STAT_EXACT    3   /* exact energy evaluations */
#define TTC_MC_STAT_OUTSIDE  4   /* proposals outside the WL window */
#define TTC_MC_STAT_PIVOT    5   /* pivot trials (the rest are displacements) */
#define TTC_MC_STAT_PIVOT_ACCEPT 6
#define TTC_MC_NSTAT 7

int  ttc_mc_create(const ttc_mc_params *p, ttc_mc **out);
int  ttc_mc_destroy(ttc_mc *m);

/* Walker w: n beads, positions (n*3), native contacts (a<b, b-a>=3) with
 * native distances, WL window [emin, emin + nbins*ebin).  At most
 * TTC_MC_MAXNAT native contacts per bead.                                 */
#define TTC_MC_MAXNAT 32
int  ttc_mc_set_walker(ttc_mc *m, int w, int n, const double *pos, int nnat, const int *na,
                       const int *nb, const double *nr0, double emin);
int  ttc_mc_get_walker(ttc_mc *m, int w, double *pos, double *energy);

/* Exact energy of every walker (bonds + native 12-10 + non-native
 * repulsion, pairs |i-j| >= 3), FP64: e[nwalk].                          */
int  ttc_mc_energy(ttc_mc *m, double *e);

/* Range energy change of walker w when beads [lb,ub) move to newpos
 * ((ub-lb)*3): sum over bonds and pairs touching [lb,ub) of
 * E(new) - E(old), as forces_mc(lb,ub,1) - forces_mc(lb,ub,0).          */
int  ttc_mc_delta_energy(ttc_mc *m, int w, int lb, int ub, const double *newpos, double *de);

/* ttc_mc_delta_energy without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The new coordinates and the
 * result pass through a fixed scratch array of the library, one per GPU
 * (TTC_MC_NOALLOC_MAXBEADS x 3 doubles + 256 B, held until the process
 * ends), instead of per-call allocations; ub - lb above the limit is
 * TTC_ERR_ARG.  Calls that overlap on one GPU (several host threads, or
 * contexts on different streams) share that array and race.              */
#define TTC_MC_NOALLOC_MAXBEADS 1024
int  ttc_mc_delta_energy_noalloc(ttc_mc *m, int w, int lb, int ub, const double *newpos, double *de);

/* Random streams of walker w: 64 xoshiro256** states of 4 words each,
 * state[4 s .. 4 s + 3] for stream s (ttc_mc_create seeds all walkers'
 * streams, walker-major, with consecutive splitmix64 outputs of seed).  The
 * counterpart of the ran2 record of the reference's restart files
 * (checkpoint.c:304-307).  Trial t of a ttc_mc_run call takes the 8 numbers
 * stream t mod 64 draws for batch t / 64 (a batch draws 8 from every stream,
 * at the start of the call and after every 64th trial but the last):
 * u0 < pivot_frac selects a pivot (bead u1 n, axis u2 3, side u3 2,
 * clockwise unless u4 2 >= 1, angle u5 pi theta_max), else a displacement
 * (bead u1 n by (u2 - 0.5, u3 - 0.5, u4 - 0.5) disp_max); u6 decides the
 * first stage, u7 the acceptance.  An all-zero state is TTC_ERR_ARG.     */
int  ttc_mc_get_streams(ttc_mc *m, int w, uint64_t *state);
int  ttc_mc_set_streams(ttc_mc *m, int w, const uint64_t *state);

/* ntrials Monte Carlo trials per walker.  mode 0: plain Wang-Landau
 * acceptance on the exact energy (ln g interpolated at the old and new
 * energies as dos_interpolate.c:107-124 does); mode 1: delayed acceptance, a surrogate
 * energy change from the network screens each proposal and the exact
 * energy makes the second-stage decision, so the chain keeps the same
 * stationary distribution.  lnf is added to ln g of the visited bin after
 * every trial (0 samples the fixed current g).  log != 0 appends
 * (features, exact dE) of every exactly evaluated move to the log buffer. */
int  ttc_mc_run(ttc_mc *m, int ntrials, int mode, double lnf, int log);

/* Device copy of the sampling state (walker positions and energies, ln g,
 * histograms, random streams): restore = 0 takes it, 1 puts it back.  The
 * in-memory counterpart of the reference's restart files for MC runs
 * (checkpoint.c:146-917); counters, network and log are not part of it.  */
int  ttc_mc_checkpoint(ttc_mc *m, int restore);

/* swap_mpi.c:24-188 (a DOS run with one energy window per MPI rank) on the
 * batch: walkers are the windows [emin, emin + nbins*ebin), and each pair
 * (w, w+1) with w = parity mod 2 exchanges configurations when each energy
 * lies inside the other's window (swap_mpi.c:147-149, the
 * msg_swap_request / msg_swap_accept exchange of defines.h:1846-1874
 * becomes device reads of the partner's state).  Every walker of a pair
 * then counts the attempt in the bin of its energy after it: h += 1,
 * ln g += lnf (swap_mpi.c:78-80, 131-133, 175-177).  The pairs must hold
 * the same chain (bead count and native contacts), else TTC_ERR_ARG.
 * accepted (optional, nwalk ints): 1 for both walkers of a swapped pair.  */
int  ttc_mc_swap_windows(ttc_mc *m, int parity, double lnf, int *accepted);
/* swap_mpi.c:190-296 merge_gofe: ln g of walker w (left window) and w + 1
 * (right) made to agree on their overlap (averaged slopes integrated from
 * the bin before the overlap, the right window's upper bins shifted to join,
 * histograms raised by (int)(change / lnf) floored at zero).  The windows
 * share the bin grid and overlap by 2 .. nbins-1 bins; the caller decides
 * when (the reference merges when mod_f < MERGE_F on its swap schedule).  */
int  ttc_mc_merge_windows(ttc_mc *m, int w, double lnf);

int  ttc_mc_get_stats(ttc_mc *m, int64_t *stats /* nwalk * TTC_MC_NSTAT */);
int  ttc_mc_reset_stats(ttc_mc *m);
int  ttc_mc_get_dos(ttc_mc *m, int w, double *lng, int64_t *hist);
int  ttc_mc_set_dos(ttc_mc *m, int w, const double *lng);
int  ttc_mc_reset_hist(ttc_mc *m);
/* copies up to max logged moves: features (max*NFEAT), exact dE (max) */
int  ttc_mc_get_log(ttc_mc *m, int max, float *features, double *de, int *count);
int  ttc_mc_clear_log(ttc_mc *m);
/* Device time of the last ttc_mc_run, microseconds. */
int  ttc_mc_last_time(ttc_mc *m, double *usec);

/* ---- surrogate network: in=NFEAT -> HID tanh -> HID tanh -> 1 ----------
 * Parameters packed W1[HID*in], b1[HID], W2[HID*HID], b2[HID], W3[HID], b3.
 * Inputs are rounded to bfloat16, all arithmetic accumulates in FP32.     */
int  ttc_mlp_nparams(void);
int  ttc_mc_set_mlp(ttc_mc *m, const float *params);
int  ttc_mc_get_mlp(ttc_mc *m, float *params);
/* forward pass of the current network on nsamp feature rows -> out */
int  ttc_mlp_forward(ttc_mc *m, int nsamp, const float *x, float *out);
/* loss 0.5*mean((f(x)-y)^2) and its gradient (FP32 accumulation) */
int  ttc_mlp_loss_grad(ttc_mc *m, int nsamp, const float *x, const float *y, double *loss, float *grad);
/* Adam training on the logged moves: targets y = asinh(dE / escale) of the
 * antisymmetric output g(old,new) - g(new,old); nepoch passes, minibatch
 * batch, learning rate lr.  Returns the final mean loss.                  */
int  ttc_mlp_train(ttc_mc *m, int nepoch, int batch, float lr, double escale, double *loss);

#ifdef __cplusplus
}
#endif
#endif
