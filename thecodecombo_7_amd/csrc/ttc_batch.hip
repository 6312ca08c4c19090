// ttc_batch.hip -- many uneven-size replicas advanced together on one GPU,
// with temperature replica exchange between them.
//
// The reference runs each replica as its own MPI process (nvt_md.c with the
// Nose-Hoover chain of integrate_nhc.c, exchanges by swap_box.c:4-86 /
// swap_mpi.c).  Here every replica is a device context (its own list, PME
// grid sized for its box, thermostat) and one call advances all of them:
// the NHC step of all replicas is captured as a HIP graph, in one of three
// launch modes:
//   mode 2: batched launches -- every kernel of the step is launched once for
//           the whole batch, its grid the concatenation of every replica's
//           own grid (a block finds its replica and its block index in a
//           device map and the replica's arguments in a device array, so it
//           computes exactly what the replica's own launch would); about 20
//           launches per step for any number of replicas;
//   mode 0: one graph whose branches are the replicas' own step graphs (one
//           launch per `spl` steps, independent branches free to overlap);
//   mode 1: one graph per replica launched on the replica's own stream.
// The launch mode and the steps per launch are chosen by ttc_batch_autotune
// from measured step times.
//
// Exchanges: replicas holding neighbouring temperatures T_t < T_t+1 are
// swapped with probability min(1, exp[(1/kT_t - 1/kT_t+1)(E_i - E_j)])
// (swap_box.c:7-11, E = en[k].potens).  The replicas differ in size, so the
// temperatures (thermostat targets) move between replicas instead of the
// configurations, and each replica's velocities are rescaled by
// sqrt(T_new/T_old) -- the same kinetic-energy matching swap_box.c:33-43
// applies to the swapped configurations.
#include "ttc_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>

int ttc_prepare_nhc(ttc_ctx *c, int max_rows, bool half);
void ttc_launch_nhc_step(ttc_ctx *c, bool energy);
extern "C" int ttc_set_nhc_temperature(ttc_ctx *c, double T, double Q);
extern "C" int ttc_tune(ttc_ctx *c, int *pair_waves, double *usec3);
extern "C" int ttc_scale_velocities(ttc_ctx *c, double factor);

// batched stages of the NHC step (kernel files)
int ttc_bstage_kinet(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_nhc(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_vscale(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_kick(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_reduce(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_drift(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_spread(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_fft(const std::vector<ttc_ctx *> &r, TtcStage *st /* [5] */);
int ttc_bstage_gather(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_rebuild(const std::vector<ttc_ctx *> &r, TtcStage *st /* [3] */);
int ttc_bstage_pair(const std::vector<ttc_ctx *> &r, TtcStage *st);
int ttc_bstage_angles(const std::vector<ttc_ctx *> &r, TtcStage *st);
void ttc_blaunch_kinet(const TtcStage &st, hipStream_t s);
void ttc_blaunch_nhc(const TtcStage &st, hipStream_t s);
void ttc_blaunch_vscale(const TtcStage &st, hipStream_t s);
void ttc_blaunch_kick(const TtcStage &st, bool energy, hipStream_t s);
void ttc_blaunch_reduce(const TtcStage &st, hipStream_t s);
void ttc_blaunch_drift(const TtcStage &st, hipStream_t s);
void ttc_blaunch_spread(const TtcStage &st, hipStream_t s);
void ttc_blaunch_fft(const TtcStage *st, bool energy, hipStream_t s);
void ttc_blaunch_gather(const TtcStage &st, bool energy, hipStream_t s);
void ttc_blaunch_rebuild(const TtcStage *st, hipStream_t s);
void ttc_blaunch_pair(const TtcStage &st, bool energy, hipStream_t s);
void ttc_blaunch_angles(const TtcStage &st, bool energy, hipStream_t s);

int ttc_stage_upload(TtcStage *st, const void *packs, size_t pack_bytes, int nrep, const std::vector<int> &blocks,
                     size_t shm) {
    ttc_stage_free(st);
    std::vector<TtcBmap> map;
    for (int i = 0; i < nrep; i++)
        for (int k = 0; k < blocks[i]; k++) map.push_back(TtcBmap{i, k});
    TTC_HIP(hipMalloc(&st->pk, pack_bytes * nrep));
    TTC_HIP(hipMalloc(&st->map, sizeof(TtcBmap) * std::max<size_t>(1, map.size())));
    TTC_HIP(hipMemcpy(st->pk, packs, pack_bytes * nrep, hipMemcpyHostToDevice));
    if (!map.empty()) TTC_HIP(hipMemcpy(st->map, map.data(), sizeof(TtcBmap) * map.size(), hipMemcpyHostToDevice));
    st->nblk = (int)map.size();
    st->shm = shm;
    return TTC_OK;
}

void ttc_stage_free(TtcStage *st) {
    if (st->pk) hipFree(st->pk);
    if (st->map) hipFree(st->map);
    *st = TtcStage{};
}

// stages of the batched NHC step
enum { BS_KINET = 0, BS_NHC, BS_VSCALE, BS_DRIFT, BS_SPREAD, BS_FFT, BS_REBUILD = BS_FFT + 5, BS_GATHER = BS_REBUILD + 3,
       BS_PAIR, BS_KICK, BS_REDUCE, BS_ANGLES, BS_N };

struct ttc_batch {
    std::vector<ttc_ctx *> r;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    int mode = 0;
    int spl = 1;
    hipGraphExec_t g_plain = nullptr, g_energy = nullptr;   // mode 0
    std::vector<hipGraphExec_t> rp, re;                      // mode 1, per replica
    hipStream_t bside = nullptr;                             // mode 2: the batch's side stream
    hipEvent_t bev[3] = {nullptr, nullptr, nullptr};
    TtcStage st[BS_N];                                       // mode 2 stages
    bool st_valid = false;
    std::vector<unsigned> gen;                               // context generations at capture
    int cap_mode = -1, cap_spl = 0;
    std::vector<int> order;                                  // order[t]: replica at temperature rank t
    uint64_t rng[4] = {0, 0, 0, 0};
    long long attempts = 0, accepts = 0;
    double ms_per_step = 0.0;
};

namespace {

#define BCHECK(cond, code, msg)          \
    do {                                 \
        if (!(cond)) {                   \
            ttc_set_error(msg);          \
            return code;                 \
        }                                \
    } while (0)

uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), uniform double in [0, 1)
double uniform(uint64_t *s) {
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    uint64_t out = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

void drop_graphs(ttc_batch *b) {
    if (b->g_plain) hipGraphExecDestroy(b->g_plain);
    if (b->g_energy) hipGraphExecDestroy(b->g_energy);
    b->g_plain = b->g_energy = nullptr;
    for (auto *v : {&b->rp, &b->re})
        for (auto &g : *v)
            if (g) { hipGraphExecDestroy(g); g = nullptr; }
    b->cap_mode = -1;
}

// one graph of `spl` NHC steps of every replica, one branch per replica
int capture_all(ttc_batch *b, bool energy, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    TTC_HIP(hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipEventRecord(b->fork, b->main);
    for (size_t i = 0; i < b->r.size() && e == hipSuccess; i++) {
        ttc_ctx *c = b->r[i];
        e = hipStreamWaitEvent(c->stream, b->fork, 0);
        for (int s = 0; s < b->spl && e == hipSuccess; s++) {
            ttc_launch_nhc_step(c, energy && s == b->spl - 1);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(b->join[i], c->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->main, b->join[i], 0);
    }
    hipError_t e2 = hipStreamEndCapture(b->main, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        ttc_set_error(std::string("ttc_batch: capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        return TTC_ERR_HIP;
    }
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    TTC_HIP(e);
    return TTC_OK;
}

// `spl` NHC steps of one replica on its own stream
int capture_one(ttc_batch *b, ttc_ctx *c, bool energy, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    TTC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < b->spl && e == hipSuccess; s++) {
        ttc_launch_nhc_step(c, energy && s == b->spl - 1);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        ttc_set_error(std::string("ttc_batch: capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        return TTC_ERR_HIP;
    }
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    TTC_HIP(e);
    return TTC_OK;
}

void drop_stages(ttc_batch *b) {
    for (TtcStage &x : b->st) ttc_stage_free(&x);
    b->st_valid = false;
}

int build_stages(ttc_batch *b) {
    drop_stages(b);
    int r = ttc_bstage_kinet(b->r, &b->st[BS_KINET]);
    if (!r) r = ttc_bstage_nhc(b->r, &b->st[BS_NHC]);
    if (!r) r = ttc_bstage_vscale(b->r, &b->st[BS_VSCALE]);
    if (!r) r = ttc_bstage_drift(b->r, &b->st[BS_DRIFT]);
    if (!r) r = ttc_bstage_spread(b->r, &b->st[BS_SPREAD]);
    if (!r) r = ttc_bstage_fft(b->r, &b->st[BS_FFT]);
    if (!r) r = ttc_bstage_rebuild(b->r, &b->st[BS_REBUILD]);
    if (!r) r = ttc_bstage_gather(b->r, &b->st[BS_GATHER]);
    if (!r) r = ttc_bstage_pair(b->r, &b->st[BS_PAIR]);
    if (!r) r = ttc_bstage_kick(b->r, &b->st[BS_KICK]);
    if (!r) r = ttc_bstage_reduce(b->r, &b->st[BS_REDUCE]);
    if (!r) r = ttc_bstage_angles(b->r, &b->st[BS_ANGLES]);
    if (r) { drop_stages(b); return r; }
    b->st_valid = true;
    return TTC_OK;
}

// one graph of `spl` NHC steps of the whole batch as batched launches, the
// same sequence as ttc_launch_nhc_step: thermostat half, drift, spread and
// FFTs (side stream) beside the conditional rebuild, gather (side) beside
// the pair kernel, kick, thermostat half (+ energies)
int capture_batched(ttc_batch *b, bool energy, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    const TtcStage *st = b->st;
    hipStream_t m = b->main, sd = b->bside;
    TTC_HIP(hipStreamBeginCapture(m, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < b->spl && e == hipSuccess; s++) {
        bool en = energy && s == b->spl - 1;
        ttc_blaunch_kinet(st[BS_KINET], m);
        ttc_blaunch_nhc(st[BS_NHC], m);
        ttc_blaunch_vscale(st[BS_VSCALE], m);
        ttc_blaunch_drift(st[BS_DRIFT], m);
        e = hipEventRecord(b->bev[0], m);
        if (e == hipSuccess) e = hipStreamWaitEvent(sd, b->bev[0], 0);
        ttc_blaunch_spread(st[BS_SPREAD], sd);
        ttc_blaunch_fft(&st[BS_FFT], en, sd);
        ttc_blaunch_rebuild(&st[BS_REBUILD], m);
        if (e == hipSuccess) e = hipEventRecord(b->bev[1], m);
        if (e == hipSuccess) e = hipStreamWaitEvent(sd, b->bev[1], 0);
        ttc_blaunch_angles(st[BS_ANGLES], en, sd);   // replicas with angle-type terms
        ttc_blaunch_gather(st[BS_GATHER], en, sd);
        ttc_blaunch_pair(st[BS_PAIR], en, m);
        if (e == hipSuccess) e = hipEventRecord(b->bev[2], sd);
        if (e == hipSuccess) e = hipStreamWaitEvent(m, b->bev[2], 0);
        ttc_blaunch_kick(st[BS_KICK], true, m);   // kinetic partials for the second thermostat half
        ttc_blaunch_nhc(st[BS_NHC], m);
        ttc_blaunch_vscale(st[BS_VSCALE], m);
        if (en) {
            ttc_blaunch_kinet(st[BS_KINET], m);
            ttc_blaunch_reduce(st[BS_REDUCE], m);
        }
        if (e == hipSuccess) e = hipGetLastError();
    }
    hipError_t e2 = hipStreamEndCapture(m, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        ttc_set_error(std::string("ttc_batch: capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        return TTC_ERR_HIP;
    }
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    TTC_HIP(e);
    return TTC_OK;
}

int sync_all(ttc_batch *b) {
    TTC_HIP(hipStreamSynchronize(b->main));
    if (b->bside) TTC_HIP(hipStreamSynchronize(b->bside));
    for (ttc_ctx *c : b->r) TTC_HIP(hipStreamSynchronize(c->stream));
    return TTC_OK;
}

int ensure_graphs(ttc_batch *b) {
    bool stale = b->cap_mode != b->mode || b->cap_spl != b->spl;
    for (size_t i = 0; i < b->r.size(); i++) stale = stale || b->gen[i] != b->r[i]->gen;
    if (!stale) return TTC_OK;
    drop_graphs(b);
    int r = sync_all(b);   // capture starts from idle streams
    if (r) return r;
    if (b->mode == 2) {
        r = build_stages(b);
        if (!r) r = capture_batched(b, false, &b->g_plain);
        if (!r) r = capture_batched(b, true, &b->g_energy);
    } else if (b->mode == 0) {
        r = capture_all(b, false, &b->g_plain);
        if (!r) r = capture_all(b, true, &b->g_energy);
    } else {
        b->rp.assign(b->r.size(), nullptr);
        b->re.assign(b->r.size(), nullptr);
        for (size_t i = 0; i < b->r.size() && !r; i++) {
            r = capture_one(b, b->r[i], false, &b->rp[i]);
            if (!r) r = capture_one(b, b->r[i], true, &b->re[i]);
        }
    }
    if (r) { drop_graphs(b); return r; }
    for (size_t i = 0; i < b->r.size(); i++) b->gen[i] = b->r[i]->gen;
    b->cap_mode = b->mode;
    b->cap_spl = b->spl;
    return TTC_OK;
}

}  // namespace

extern "C" int ttc_batch_create(ttc_ctx *const *ctx, int nrep, uint64_t seed, ttc_batch **out) {
    BCHECK(ctx && out && nrep >= 1, TTC_ERR_ARG, "ttc_batch_create: bad argument");
    for (int i = 0; i < nrep; i++)
        BCHECK(ctx[i] && ctx[i]->topo && ctx[i]->nhc_M > 0, TTC_ERR_ARG,
               "ttc_batch_create: every replica needs a topology and a thermostat (ttc_set_nhc)");
    ttc_batch *b = new ttc_batch();
    b->r.assign(ctx, ctx + nrep);
    b->gen.assign(nrep, ~0u);
    b->join.assign(nrep, nullptr);
    hipError_t e = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->bside, hipStreamNonBlocking);
    for (int k = 0; k < 3 && e == hipSuccess; k++) e = hipEventCreateWithFlags(&b->bev[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int i = 0; i < nrep && e == hipSuccess; i++) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        ttc_set_error(std::string("ttc_batch_create: ") + hipGetErrorString(e));
        delete b;
        return TTC_ERR_HIP;
    }
    b->order.resize(nrep);
    for (int i = 0; i < nrep; i++) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int x, int y) { return b->r[x]->nhc_T < b->r[y]->nhc_T; });
    uint64_t sm = seed;
    for (int k = 0; k < 4; k++) b->rng[k] = splitmix(sm);
    *out = b;
    return TTC_OK;
}

extern "C" int ttc_batch_destroy(ttc_batch *b) {
    if (!b) return TTC_OK;
    drop_graphs(b);
    drop_stages(b);
    for (hipEvent_t e : b->join)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : b->bev)
        if (e) hipEventDestroy(e);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->bside) hipStreamDestroy(b->bside);
    if (b->main) hipStreamDestroy(b->main);
    delete b;
    return TTC_OK;
}

extern "C" int ttc_batch_set_launch(ttc_batch *b, int mode, int steps_per_launch) {
    BCHECK(b && mode >= 0 && mode <= 2 && steps_per_launch >= 1 && steps_per_launch <= 1000, TTC_ERR_ARG,
           "ttc_batch_set_launch: mode 0/1/2, 1 <= steps_per_launch <= 1000");
    b->mode = mode;
    b->spl = steps_per_launch;
    return TTC_OK;
}

// half lists for the NHC steps when every replica takes them (the batched
// pair stage runs one kernel variant for all)
static bool batch_half(const ttc_batch *b) {
    for (const ttc_ctx *c : b->r)
        if (!ttc_half_enabled(c)) return false;
    return true;
}

extern "C" int ttc_batch_md_nhc(ttc_batch *b, int nsteps, double *potens, int *rebuilds) {
    BCHECK(b && nsteps >= 0 && nsteps % b->spl == 0, TTC_ERR_ARG,
           "ttc_batch_md_nhc: nsteps must be a multiple of the steps per launch");
    size_t R = b->r.size();
    for (ttc_ctx *c : b->r) {
        int r = ttc_prepare_nhc(c, 1, batch_half(b));
        if (r) return r;
    }
    int r = ensure_graphs(b);
    if (r) return r;
    for (ttc_ctx *c : b->r) {
        TTC_HIP(hipMemsetAsync(c->status, 0, sizeof(int) * ST_NSTAT, c->stream));
        ttc_launch_nlcheck(c);
    }
    r = sync_all(b);
    if (r) return r;
    auto t0 = std::chrono::steady_clock::now();
    int nl = nsteps / b->spl;
    for (int l = 0; l < nl; l++) {
        bool en = l == nl - 1;
        if (b->mode == 0 || b->mode == 2) {
            TTC_HIP(hipGraphLaunch(en ? b->g_energy : b->g_plain, b->main));
        } else {
            for (size_t i = 0; i < R; i++) TTC_HIP(hipGraphLaunch(en ? b->re[i] : b->rp[i], b->r[i]->stream));
        }
    }
    r = sync_all(b);
    if (r) return r;
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    b->ms_per_step = nsteps > 0 ? ms / nsteps : 0.0;
    for (size_t i = 0; i < R; i++) {
        ttc_ctx *c = b->r[i];
        TTC_HIP(hipMemcpy(c->h_status, c->status, sizeof(int) * ST_NSTAT, hipMemcpyDeviceToHost));
        if (c->h_status[ST_OVERFLOW]) {
            ttc_set_error("ttc_batch_md_nhc: neighbour list overflow in replica " + std::to_string(i) +
                          " (nblist.c:185-196; flags " + std::to_string(c->h_status[ST_OVERFLOW]) +
                          ", longest row " + std::to_string(c->h_status[ST_MAXROW]) + " of " +
                          std::to_string(c->cap) + ", window " + std::to_string(c->h_status[ST_WMAX]) + " of " +
                          std::to_string(c->wcap) + "); call ttc_nblist on it and run again");
            return TTC_ERR_LIST_OVERFLOW;
        }
        if (rebuilds) rebuilds[i] = c->h_status[ST_NREBUILD];
        if (c->h_status[ST_NREBUILD] > 0) c->list_pruned = true;
        if (potens && nsteps > 0) {
            double row[ROW_LEN];
            TTC_HIP(hipMemcpy(row, c->rows, sizeof row, hipMemcpyDeviceToHost));
            potens[i] = row[2] + row[4];   // calcvalue.c:40: bond + nbonds
        }
        c->f4_valid = true;
    }
    return TTC_OK;
}

extern "C" int ttc_batch_exchange(ttc_batch *b, const double *potens, int parity, int *accepted) {
    BCHECK(b && potens && (parity == 0 || parity == 1), TTC_ERR_ARG, "ttc_batch_exchange: bad argument");
    int acc = 0;
    const double rg = TTC_KB * TTC_NA * 0.001;   // sim.kT = RG T .001 (read_sim.c:563), kJ/mol
    for (size_t t = parity; t + 1 < b->order.size(); t += 2) {
        int i = b->order[t], j = b->order[t + 1];
        ttc_ctx *ci = b->r[i], *cj = b->r[j];
        double Ti = ci->nhc_T, Tj = cj->nhc_T;
        double delta_beta = 1.0 / (rg * Ti) - 1.0 / (rg * Tj);
        double crit = exp(delta_beta * (potens[i] - potens[j]));
        b->attempts++;
        if (crit > uniform(b->rng)) {
            int r = ttc_set_nhc_temperature(ci, Tj, ci->nhc_Q * Tj / Ti);
            if (!r) r = ttc_set_nhc_temperature(cj, Ti, cj->nhc_Q * Ti / Tj);
            if (!r) r = ttc_scale_velocities(ci, sqrt(Tj / Ti));
            if (!r) r = ttc_scale_velocities(cj, sqrt(Ti / Tj));
            if (r) return r;
            std::swap(b->order[t], b->order[t + 1]);
            b->accepts++;
            acc++;
        }
    }
    if (accepted) *accepted = acc;
    return TTC_OK;
}

extern "C" int ttc_batch_state(ttc_batch *b, int *order, double *temps, double *ms_per_step, long long *attempts,
                               long long *accepts) {
    BCHECK(b, TTC_ERR_ARG, "ttc_batch_state: bad argument");
    for (size_t t = 0; t < b->r.size(); t++) {
        if (order) order[t] = b->order[t];
        if (temps) temps[t] = b->r[t]->nhc_T;
    }
    if (ms_per_step) *ms_per_step = b->ms_per_step;
    if (attempts) *attempts = b->attempts;
    if (accepts) *accepts = b->accepts;
    return TTC_OK;
}

// First tunes each replica's pair-kernel shape (ttc_tune; the batched
// launches of mode 2 use 8 waves per group for every replica), then tries
// every (launch mode, steps per launch) candidate for `trial_steps` steps
// (the steps are real dynamics: they count as equilibration) and keeps the
// fastest; the measured ms/step of each candidate is returned in
// timings[12] if given.
extern "C" int ttc_batch_autotune(ttc_batch *b, int trial_steps, int *best_mode, int *best_spl, double *timings) {
    BCHECK(b && trial_steps >= 1, TTC_ERR_ARG, "ttc_batch_autotune: bad argument");
    for (ttc_ctx *c : b->r) {
        int r = ttc_prepare_nhc(c, 1, batch_half(b));
        if (!r) r = ttc_tune(c, nullptr, nullptr);
        if (r) return r;
    }
    const int spls[] = {1, 5, 10, 25};
    double best = 1e300;
    int bm = 0, bs = 1, q = 0;
    for (int mode = 0; mode < 3; mode++)
        for (int spl : spls) {
            b->mode = mode;
            b->spl = spl;
            int steps = std::max(spl, (trial_steps / spl) * spl);
            int r = ttc_batch_md_nhc(b, spl, nullptr, nullptr);   // capture + warm
            if (!r) r = ttc_batch_md_nhc(b, steps, nullptr, nullptr);
            if (r) return r;
            if (timings) timings[q] = b->ms_per_step;
            q++;
            if (b->ms_per_step < best) { best = b->ms_per_step; bm = mode; bs = spl; }
        }
    b->mode = bm;
    b->spl = bs;
    if (best_mode) *best_mode = bm;
    if (best_spl) *best_spl = bs;
    return TTC_OK;
}
