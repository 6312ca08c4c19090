// ttc_api.hip -- the C ABI of libttc.so (include/ttc.h): allocation,
// topology upload, host <-> device copies, the reference-named operations
// and the graph-captured velocity-Verlet loop.
#include "ttc_common.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <map>

int ttc_pair_half_attrs(ttc_ctx *c);
int ttc_build_grp_attrs(ttc_ctx *c);
void ttc_launch_snap(ttc_ctx *c);
void ttc_launch_posq(ttc_ctx *c);
void ttc_launch_pair_only(ttc_ctx *c);

static thread_local std::string g_err;

// a pair of timing events destroyed on every return path
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~EventPair() {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    }
    hipError_t create() {
        hipError_t r = hipEventCreate(&e0);
        return r != hipSuccess ? r : hipEventCreate(&e1);
    }
};

void ttc_set_error(const std::string &msg) { g_err = msg; }
extern "C" const char *ttc_last_error(void) { return g_err.c_str(); }

// ewald.c:817-835 (Numerical Recipes erfcc) in FP64 for the host-side
// constant erfc(kappa rc)/rc.
double ttc_nr_erfc_host(double x) {
    double z = fabs(x);
    double t = 1.0 / (1.0 + 0.5 * z);
    double v = t * exp(-z * z + -1.26551223 +
               t * (1.00002368 + t * (0.37409196 + t * (0.09678418 +
               t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 +
               t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
    return x < 0 ? 2.0 - v : v;
}

#define CHECK(cond, code, msg)            \
    do {                                  \
        if (!(cond)) {                    \
            ttc_set_error(msg);           \
            return code;                  \
        }                                 \
    } while (0)

template <typename T>
static int dalloc(T **p, size_t n) {
    if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
        ttc_set_error("hipMalloc failed");
        return TTC_ERR_ALLOC;
    }
    return TTC_OK;
}

#define ALLOC(p, n)                      \
    do {                                 \
        int r_ = dalloc(&(p), (n));      \
        if (r_) return r_;               \
    } while (0)

static int list_capacity(const ttc_ctx *c, double cutoff) {
    double vol = c->p.lx * c->p.ly * c->p.lz;
    double rho = c->n / vol;
    double expect = 4.0 / 3.0 * TTC_PI * cutoff * cutoff * cutoff * rho;
    int cap = (int)(expect * 1.25) + 64;
    cap = (cap + 63) & ~63;
    return std::min(cap, std::max(64, ((c->n + 63) / 64) * 64));
}

extern "C" int ttc_create(const ttc_params *p, ttc_ctx **out) {
    CHECK(p && out, TTC_ERR_ARG, "ttc_create: null argument");
    CHECK(p->n > 0 && p->n < TTC_MAX_SITES, TTC_ERR_ARG, "ttc_create: 1 .. 2^24 - 1 sites supported");
    CHECK(p->lx > 0 && p->ly > 0 && p->lz > 0, TTC_ERR_ARG, "ttc_create: bad box");
    CHECK(p->rc > 0 && p->rclist >= p->rc, TTC_ERR_ARG, "ttc_create: need rclist >= rc > 0");
    CHECK(p->order == 4 || p->order == 6, TTC_ERR_ARG, "ttc_create: spline order must be 4 or 6");
    CHECK(p->gx >= p->order && p->gy >= p->order && p->gz >= p->order, TTC_ERR_ARG,
          "ttc_create: PME grid smaller than the spline order");
    CHECK(p->gx <= 1024 && p->gy <= 1024 && p->gz <= 1024, TTC_ERR_ARG, "ttc_create: PME grid too large");
    CHECK(p->ntypes > 0 && p->ntypes <= TTC_MAX_TYPES, TTC_ERR_ARG, "ttc_create: 1..64 site types supported");
    double lmin = std::min(p->lx, std::min(p->ly, p->lz));
    CHECK(2.0 * p->rc <= lmin, TTC_ERR_ARG, "ttc_create: cutoff exceeds half the box");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        ttc_set_error("ttc_create: no HIP device");
        return TTC_ERR_HIP;
    }
    ttc_ctx *c = new ttc_ctx();
    c->p = *p;
    c->n = p->n;
    c->G = (p->n + 63) / 64;
    c->ntypes = p->ntypes;
    c->order = p->order;
    TTC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    TTC_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    for (int k = 0; k < 3; k++) TTC_HIP(hipEventCreateWithFlags(&c->sev[k], hipEventDisableTiming));
    int n = c->n;
    ALLOC(c->pos4, n); ALLOC(c->vel4, n); ALLOC(c->f4, n); ALLOC(c->anchor, n);
    ALLOC(c->pos4b, n); ALLOC(c->vel4b, n); ALLOC(c->f4b, n); ALLOC(c->posfb, n);
    ALLOC(c->type, n); ALLOC(c->typeb, n); ALLOC(c->orig, n); ALLOC(c->origb, n); ALLOC(c->inv, n);
    ALLOC(c->fpair, n); ALLOC(c->g4, n); ALLOC(c->snap, n); ALLOC(c->g4o, n);
    ALLOC(c->fs, n); ALLOC(c->fl, n); ALLOC(c->fsb, n); ALLOC(c->flb, n);
    TTC_HIP(hipMemset(c->fs, 0, sizeof(double4) * n));
    TTC_HIP(hipMemset(c->fl, 0, sizeof(double4) * n));
    ALLOC(c->nhc, TTC_NHC_LEN);
    TTC_HIP(hipMemset(c->nhc, 0, sizeof(double) * TTC_NHC_LEN));
    TTC_HIP(hipMemset(c->pos4, 0, sizeof(double4) * n));
    TTC_HIP(hipMemset(c->f4, 0, sizeof(double4) * n));
    TTC_HIP(hipMemset(c->fpair, 0, sizeof(double4) * n));
    TTC_HIP(hipMemset(c->g4, 0, sizeof(double4) * n));
    std::vector<int> ident(n);
    for (int i = 0; i < n; i++) ident[i] = i;
    TTC_HIP(hipMemcpy(c->orig, ident.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    TTC_HIP(hipMemcpy(c->inv, ident.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    std::vector<double4> v0(n, make_double4(0, 0, 0, 100.0));
    TTC_HIP(hipMemcpy(c->vel4, v0.data(), sizeof(double4) * n, hipMemcpyHostToDevice));

    // cells at least (widest list cutoff)/2 wide
    double rmax = p->ewald_lists ? std::max(p->rclist, p->rclist_ew) : p->rclist;
    c->ncx = std::max(1, (int)(p->lx / (0.5 * rmax)));
    c->ncy = std::max(1, (int)(p->ly / (0.5 * rmax)));
    c->ncz = std::max(1, (int)(p->lz / (0.5 * rmax)));
    c->ncell = c->ncx * c->ncy * c->ncz;
    ALLOC(c->cell_of, n); ALLOC(c->cell_cnt, c->ncell); ALLOC(c->cell_start, c->ncell);
    ALLOC(c->cell_acc, c->ncell);
    TTC_HIP(hipMemset(c->cell_acc, 0, sizeof(int) * c->ncell));
    // the MD spread's copies start empty (no sites in any cell)
    ALLOC(c->spos, n); ALLOC(c->scs, c->ncell); ALLOC(c->scc, c->ncell);
    TTC_HIP(hipMemset(c->spos, 0, sizeof(double4) * n));
    TTC_HIP(hipMemset(c->scs, 0, sizeof(int) * c->ncell));
    TTC_HIP(hipMemset(c->scc, 0, sizeof(int) * c->ncell));
    ALLOC(c->cell_members, (size_t)c->ncell * ttc_cellcap(c));

    c->cap = list_capacity(c, ttc_list_cutoff(c));
    ALLOC(c->nbl, (size_t)c->G * c->cap * 64);
    ALLOC(c->glen, c->G);
    ALLOC(c->gcnt, c->G);
    TTC_HIP(hipMemset(c->glen, 0, sizeof(int) * c->G));
    ALLOC(c->status, ST_NSTAT);
    TTC_HIP(hipMemset(c->status, 0, sizeof(int) * ST_NSTAT));
    TTC_HIP(hipHostMalloc((void **)&c->h_status, sizeof(int) * ST_NSTAT, hipHostMallocDefault));
    memset(c->h_status, 0, sizeof(int) * ST_NSTAT);

    // half-list state: partner accumulators (cleared by every combine that
    // reads them), windows sized for this geometry
    ALLOC(c->fixj, (size_t)3 * n);
    TTC_HIP(hipMemset(c->fixj, 0, sizeof(long long) * 3 * n));
    ttc_window_sizes(c, &c->wcap, &c->span, &c->span_pre);
    ALLOC(c->posq, n);
    ALLOC(c->win, (size_t)c->G * c->wcap);
    ALLOC(c->wlen, c->G);
    TTC_HIP(hipMemset(c->wlen, 0, sizeof(int) * c->G));
    int r = ttc_pair_half_attrs(c);
    if (r) return r;
    r = ttc_build_grp_attrs(c);
    if (r) return r;
    r = ttc_pme_setup(c);
    if (r) return r;
    r = ttc_pme_attrs(c);
    if (r) return r;
    c->nblk_max = std::max(std::max(ttc_pair_blocks(c), ttc_pme_conv_blocks(c)),
                           std::max(ttc_gather_blocks(c), ttc_kick_blocks(c)));
    ALLOC(c->partials, (size_t)EV_N * c->nblk_max);
    TTC_HIP(hipMemset(c->partials, 0, sizeof(double) * EV_N * c->nblk_max));
    ALLOC(c->ev, EV_N + 4);
    double chk[4] = {p->lx, p->ly, p->lz, (p->rclist - p->rc) / 4.0};
    TTC_HIP(hipMemcpy(c->ev + EV_N, chk, sizeof(chk), hipMemcpyHostToDevice));
    TTC_HIP(hipHostMalloc((void **)&c->h_ev, sizeof(double) * (EV_N + 4), hipHostMallocDefault));
    *out = c;
    return TTC_OK;
}

static void destroy_graphs(ttc_ctx *c) {
    hipGraphExec_t *gs[] = {&c->g_vv[0][0], &c->g_vv[0][1], &c->g_vv[1][0], &c->g_vv[1][1], &c->g_vv_run, &c->g_nhc_plain,
                            &c->g_nhc_energy, &c->g_mts_plain, &c->g_mts_energy};
    for (hipGraphExec_t *g : gs) {
        if (*g) hipGraphExecDestroy(*g);
        *g = nullptr;
    }
}

extern "C" int ttc_destroy(ttc_ctx *c) {
    if (!c) return TTC_OK;
    hipStreamSynchronize(c->stream);
    destroy_graphs(c);
    void *bufs[] = {c->pos4, c->vel4, c->f4, c->anchor, c->pos4b, c->vel4b, c->f4b, c->type, c->typeb, c->orig,
                    c->origb, c->inv, c->fpair, c->g4, c->cell_of, c->cell_cnt, c->cell_start, c->cell_acc,
                    c->cell_members, c->nbl, c->glen, c->status, c->ex_off, c->ex_part, c->ex_flag, c->ljtab,
                    c->b_off, c->b_part, c->b_r0, c->b_k, c->partials, c->ev, c->rows, c->fs, c->fl, c->fsb,
                    c->flb, c->nhc, c->posfb, c->gcnt, c->snap, c->g4o, c->fixj, c->win, c->wlen, c->ang_idx,
                    c->ang_par, c->ang_tpar, c->ang_tf, c->ang_off, c->ang_ref, c->posq, c->spos, c->scs, c->scc};
    for (void *b : bufs)
        if (b) hipFree(b);
    ttc_pme_free(c);
    if (c->h_status) hipHostFree(c->h_status);
    if (c->h_ev) hipHostFree(c->h_ev);
    hipStreamSynchronize(c->side);
    for (int k = 0; k < 3; k++)
        if (c->sev[k]) hipEventDestroy(c->sev[k]);
    hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    delete c;
    return TTC_OK;
}

extern "C" double ttc_self_energy(const ttc_ctx *c) { return c ? c->e_self : 0.0; }

// device sorted order -> host original order helpers
static int fetch_orig(ttc_ctx *c, std::vector<int> &orig) {
    orig.resize(c->n);
    TTC_HIP(hipMemcpyAsync(orig.data(), c->orig, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    return TTC_OK;
}

extern "C" int ttc_set_topology(ttc_ctx *c, const int *atomid, const double *qq, const double *mass,
                                const double *lj_eps, const double *lj_sig, const double *lj14_eps,
                                const double *lj14_sig, int nx, const int *xa, const int *xb, const int *xflag,
                                int nbond, const int *ba, const int *bb, const double *req, const double *kbond) {
    CHECK(c && atomid && qq && mass && lj_eps && lj_sig && lj14_eps && lj14_sig, TTC_ERR_ARG,
          "ttc_set_topology: null argument");
    int n = c->n, nt = c->ntypes;
    for (int i = 0; i < n; i++) {
        CHECK(atomid[i] >= 0 && atomid[i] < nt, TTC_ERR_ARG, "ttc_set_topology: atomid out of range");
        CHECK(mass[i] > 0, TTC_ERR_ARG, "ttc_set_topology: non-positive mass");
    }
    // ewald.c:184-194 neutrality, ewald.c:171-178 self energy
    double charge = 0.0, q2 = 0.0;
    for (int i = 0; i < n; i++) { charge += qq[i]; q2 += qq[i] * qq[i]; }
    if (fabs(charge) > 0.000005) {
        char buf[160];
        snprintf(buf, sizeof(buf), "The charge of the system is %f: Ewald summation needs a neutral system", charge);
        ttc_set_error(buf);
        return TTC_ERR_NOT_NEUTRAL;
    }
    double fconv = (8.987551789e19 * 1.602176462e-19 * 1.602176462e-19 * 6.022045E23 * 0.001) / c->p.eps_rf;
    c->e_self = q2 * (c->p.kappa / sqrt(TTC_PI)) * fconv;

    std::vector<int> orig;
    int r = fetch_orig(c, orig);
    if (r) return r;
    std::vector<int> typ(n);
    std::vector<double4> pos(n), vel(n);
    TTC_HIP(hipMemcpy(pos.data(), c->pos4, sizeof(double4) * n, hipMemcpyDeviceToHost));
    TTC_HIP(hipMemcpy(vel.data(), c->vel4, sizeof(double4) * n, hipMemcpyDeviceToHost));
    for (int s = 0; s < n; s++) {
        int o = orig[s];
        typ[s] = atomid[o];
        pos[s].w = qq[o];
        vel[s].w = 100.0 / mass[o];
    }
    TTC_HIP(hipMemcpy(c->type, typ.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    TTC_HIP(hipMemcpy(c->pos4, pos.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    TTC_HIP(hipMemcpy(c->vel4, vel.data(), sizeof(double4) * n, hipMemcpyHostToDevice));

    // LJ tables
    std::vector<float4> lj((size_t)nt * nt);
    for (int a = 0; a < nt * nt; a++)
        lj[a] = make_float4((float)lj_eps[a], (float)(lj_sig[a] * lj_sig[a]), (float)lj14_eps[a],
                            (float)(lj14_sig[a] * lj14_sig[a]));
    if (c->ljtab) hipFree(c->ljtab);
    ALLOC(c->ljtab, (size_t)nt * nt);
    TTC_HIP(hipMemcpy(c->ljtab, lj.data(), sizeof(float4) * nt * nt, hipMemcpyHostToDevice));

    // exclusions: per-site rows, flag of the last xint entry naming the pair
    std::vector<std::map<int, int>> ex(n);
    int span = 0;
    for (int i = 0; i < nx; i++) {
        int a = xa[i], b = xb[i];
        CHECK(a >= 0 && a < n && b >= 0 && b < n && a != b, TTC_ERR_ARG, "ttc_set_topology: bad exclusion");
        ex[a][b] = xflag[i];
        ex[b][a] = xflag[i];
        span = std::max(span, std::abs(a - b));
    }
    std::vector<int> off(n + 1, 0), part, flg;
    for (int a = 0; a < n; a++) {
        off[a + 1] = off[a] + (int)ex[a].size();
        for (auto &kv : ex[a]) { part.push_back(kv.first); flg.push_back(kv.second); }
    }
    c->ex_span = span;
    c->nx = nx;
    if (c->ex_off) { hipFree(c->ex_off); hipFree(c->ex_part); hipFree(c->ex_flag); }
    ALLOC(c->ex_off, n + 1); ALLOC(c->ex_part, part.size()); ALLOC(c->ex_flag, flg.size());
    TTC_HIP(hipMemcpy(c->ex_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (!part.empty()) {
        TTC_HIP(hipMemcpy(c->ex_part, part.data(), sizeof(int) * part.size(), hipMemcpyHostToDevice));
        TTC_HIP(hipMemcpy(c->ex_flag, flg.data(), sizeof(int) * flg.size(), hipMemcpyHostToDevice));
    }

    // bonds, both directions
    std::vector<std::vector<int>> bl(n);
    for (int i = 0; i < nbond; i++) {
        CHECK(ba[i] >= 0 && ba[i] < n && bb[i] >= 0 && bb[i] < n && ba[i] != bb[i], TTC_ERR_ARG,
              "ttc_set_topology: bad bond");
        bl[ba[i]].push_back(i);
        bl[bb[i]].push_back(i);
    }
    std::vector<int> boff(n + 1, 0), bpart;
    std::vector<double> br0, bk;
    for (int a = 0; a < n; a++) {
        boff[a + 1] = boff[a] + (int)bl[a].size();
        for (int i : bl[a]) {
            bpart.push_back(ba[i] == a ? bb[i] : ba[i]);
            br0.push_back(req[i]);
            bk.push_back(kbond[i]);
        }
    }
    c->nbond = nbond;
    if (c->b_off) { hipFree(c->b_off); hipFree(c->b_part); hipFree(c->b_r0); hipFree(c->b_k); }
    ALLOC(c->b_off, n + 1); ALLOC(c->b_part, bpart.size()); ALLOC(c->b_r0, br0.size()); ALLOC(c->b_k, bk.size());
    TTC_HIP(hipMemcpy(c->b_off, boff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (!bpart.empty()) {
        TTC_HIP(hipMemcpy(c->b_part, bpart.data(), sizeof(int) * bpart.size(), hipMemcpyHostToDevice));
        TTC_HIP(hipMemcpy(c->b_r0, br0.data(), sizeof(double) * br0.size(), hipMemcpyHostToDevice));
        TTC_HIP(hipMemcpy(c->b_k, bk.data(), sizeof(double) * bk.size(), hipMemcpyHostToDevice));
    }
    c->topo = true;
    c->f4_valid = false;
    destroy_graphs(c);
    return TTC_OK;
}

static int put_xyz(ttc_ctx *c, const double *x, int64_t stride, bool velocity) {
    CHECK(c && x && stride >= 3, TTC_ERR_ARG, "ttc_set_*: bad argument");
    std::vector<int> orig;
    int r = fetch_orig(c, orig);
    if (r) return r;
    int n = c->n;
    double4 *dst = velocity ? c->vel4 : c->pos4;
    std::vector<double4> h(n);
    TTC_HIP(hipMemcpy(h.data(), dst, sizeof(double4) * n, hipMemcpyDeviceToHost));
    for (int s = 0; s < n; s++) {
        const double *src = x + (size_t)orig[s] * stride;
        h[s].x = src[0]; h[s].y = src[1]; h[s].z = src[2];
    }
    TTC_HIP(hipMemcpy(dst, h.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    if (!velocity) {
        c->f4_valid = false;
        c->mts_valid = false;
    }
    return TTC_OK;
}

static int get_xyz(ttc_ctx *c, double *x, int64_t stride, const double4 *src, bool add) {
    CHECK(c && x && stride >= 3, TTC_ERR_ARG, "ttc_get_*: bad argument");
    std::vector<int> orig;
    int r = fetch_orig(c, orig);
    if (r) return r;
    int n = c->n;
    std::vector<double4> h(n);
    TTC_HIP(hipMemcpyAsync(h.data(), src, sizeof(double4) * n, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < n; s++) {
        double *d = x + (size_t)orig[s] * stride;
        if (add) { d[0] += h[s].x; d[1] += h[s].y; d[2] += h[s].z; }
        else { d[0] = h[s].x; d[1] = h[s].y; d[2] = h[s].z; }
    }
    return TTC_OK;
}

extern "C" int ttc_set_positions(ttc_ctx *c, const double *x, int64_t stride) { return put_xyz(c, x, stride, false); }
extern "C" int ttc_get_positions(ttc_ctx *c, double *x, int64_t stride) { return get_xyz(c, x, stride, c->pos4, false); }
extern "C" int ttc_set_velocities(ttc_ctx *c, const double *v, int64_t stride) { return put_xyz(c, v, stride, true); }
extern "C" int ttc_get_velocities(ttc_ctx *c, double *v, int64_t stride) { return get_xyz(c, v, stride, c->vel4, false); }

static int read_status(ttc_ctx *c) {
    TTC_HIP(hipMemcpyAsync(c->h_status, c->status, sizeof(int) * ST_NSTAT, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    return TTC_OK;
}

static int grow_list(ttc_ctx *c, int need) {
    if (c->row_limit > 0 && need <= c->cap) {
        // a limit below the capacity reported it: lift the limit
        c->row_limit = 0;
        destroy_graphs(c);
        return TTC_OK;
    }
    hipFree(c->nbl);
    c->cap = ((int)(need * 1.2) + 127) & ~63;
    ALLOC(c->nbl, (size_t)c->G * c->cap * 64);
    c->row_limit = 0;
    destroy_graphs(c);
    c->gen++;
    return TTC_OK;
}

// wider half-list windows or builder bitmap after a build reported wmax
// partners in one group, a partner at bitmap bit spanmax - 1, or a partner
// premax sites before a group's first site
static int grow_window(ttc_ctx *c, int wmax, int spanmax, int premax) {
    if (wmax > c->wcap) {
        hipFree(c->win);
        c->wcap = ((int)(wmax * 1.2) + 63) & ~63;
        ALLOC(c->win, (size_t)c->G * c->wcap);
    }
    if (spanmax > c->span || premax > c->span_pre) {
        int below = std::max(c->span_pre, (int)(premax * 1.2));
        int above = std::max(c->span - c->span_pre, (int)((spanmax - c->span_pre) * 1.2));
        ttc_bitmap_sizes(c, below, above, &c->span, &c->span_pre);
    }
    c->win_limit = 0;
    size_t ttc_pair_half_shm(const ttc_ctx *c);
    if (ttc_pair_half_shm(c) > 150 * 1024) {
        ttc_set_error("half-list window too wide for one CU's LDS (sites far denser than the box average)");
        return TTC_ERR_LIST_OVERFLOW;
    }
    int r = ttc_pair_half_attrs(c);
    if (r) return r;
    r = ttc_build_grp_attrs(c);
    if (r) return r;
    destroy_graphs(c);
    c->gen++;
    return TTC_OK;
}

// grows whatever the last build reported too small (status already read);
// returns TTC_ERR_LIST_OVERFLOW for a cell overflow, which growing cannot fix
static int grow_after_overflow(ttc_ctx *c) {
    const int ov = c->h_status[ST_OVERFLOW];
    if (ov & 2) {
        ttc_set_error("a neighbour cell holds more sites than its capacity");
        return TTC_ERR_LIST_OVERFLOW;
    }
    if (ov & 1) {
        int r = grow_list(c, c->h_status[ST_MAXROW]);
        if (r) return r;
    }
    if (ov & 8) {
        int r = grow_window(c, c->h_status[ST_WMAX], c->h_status[ST_SPANMAX], c->h_status[ST_PREMAX]);
        if (r) return r;
    }
    return TTC_OK;
}

// clears the overflow flag and the maxima the builds report
static int clear_overflow(ttc_ctx *c) {
    TTC_HIP(hipMemsetAsync(c->status + ST_OVERFLOW, 0, sizeof(int) * 2, c->stream));
    TTC_HIP(hipMemsetAsync(c->status + ST_WMAX, 0, sizeof(int) * 3, c->stream));
    return TTC_OK;
}

// nblist.c:32: unconditional rebuild; grows the device list if a row
// did not fit (the reference would exit(5) instead)
extern "C" int ttc_nblist(ttc_ctx *c) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_nblist: topology not set");
    for (int attempt = 0; attempt < 3; attempt++) {
        TTC_HIP(hipMemsetAsync(c->status, 0, sizeof(int) * 4, c->stream));
        int r = ttc_launch_rebuild(c, false);
        if (r) return r;
        TTC_HIP(hipGetLastError());
        r = read_status(c);
        if (r) return r;
        int ov = c->h_status[ST_OVERFLOW];
        if (ov & 2) {
            ttc_set_error("ttc_nblist: a neighbour cell holds more sites than its capacity");
            return TTC_ERR_LIST_OVERFLOW;
        }
        if (ov & 1) {
            r = grow_list(c, c->h_status[ST_MAXROW]);
            if (r) return r;
            continue;
        }
        c->have_list = true;
        c->list_pruned = false;
        c->list_half = false;
        c->f4_valid = false;   // the sorted order changed
        return TTC_OK;
    }
    ttc_set_error("ttc_nblist: list capacity could not be grown");
    return TTC_ERR_LIST_OVERFLOW;
}

int ttc_launch_build_at_anchors(ttc_ctx *c);

int ttc_rebuild_full_at_anchors(ttc_ctx *c) {
    for (int attempt = 0; attempt < 3; attempt++) {
        TTC_HIP(hipMemsetAsync(c->status + ST_OVERFLOW, 0, sizeof(int) * 2, c->stream));
        int r = ttc_launch_build_at_anchors(c);
        if (r) return r;
        TTC_HIP(hipGetLastError());
        r = read_status(c);
        if (r) return r;
        if (!(c->h_status[ST_OVERFLOW] & 1)) return TTC_OK;
        r = grow_list(c, c->h_status[ST_MAXROW]);
        if (r) return r;
    }
    ttc_set_error("ttc: list capacity could not be grown");
    return TTC_ERR_LIST_OVERFLOW;
}

extern "C" int ttc_nl_check(ttc_ctx *c, int lb, int ub, int *nl_flag) {
    CHECK(c && nl_flag && lb >= 0 && ub <= c->n && lb <= ub, TTC_ERR_ARG, "ttc_nl_check: bad argument");
    CHECK(c->have_list, TTC_ERR_ARG, "ttc_nl_check: no list built");
    return ttc_nl_check_range(c, lb, ub, nl_flag);
}

extern "C" int ttc_get_nblist(ttc_ctx *c, int which, int max_nb, int *count, int *list, int *xflag) {
    CHECK(c && count && list && xflag && max_nb > 0, TTC_ERR_ARG, "ttc_get_nblist: bad argument");
    double cutoff;
    if (which == TTC_LIST_NLIST) cutoff = ttc_list_cutoff(c);
    else if (which == TTC_LIST_NLIST_EW) {
        CHECK(c->p.ewald_lists, TTC_ERR_ARG, "ttc_get_nblist: nlist_ew exists only with ewald_lists");
        cutoff = c->p.rclist_ew;
    } else {
        ttc_set_error("ttc_get_nblist: which must be TTC_LIST_NLIST or TTC_LIST_NLIST_EW");
        return TTC_ERR_ARG;
    }
    return ttc_export_list(c, cutoff, max_nb, count, list, xflag);
}

// The reference evaluates with the list of its last nblist() whatever the
// positions are now.  A list made inside an MD loop is pruned to
// rc + 2 (rclist-rc)/4 (complete while no site moved more than the DLIST
// distance); if a site moved further since, the list is rebuilt at the full
// rclist from the anchors -- the reference's list -- before evaluating.
// Kernels other than the plain pair kernel need the full format: a half
// list is replaced by the reference's list at the anchors unless allow_half.
static int need_list(ttc_ctx *c, bool allow_half = false) {
    CHECK(c && c->topo, TTC_ERR_ARG, "topology not set");
    if (!c->have_list) return ttc_nblist(c);
    bool rebuild = c->list_half && !allow_half;
    if (!rebuild && c->list_pruned) {
        int flag = 0;
        int r = ttc_nl_check_range(c, 0, c->n, &flag);
        if (r) return r;
        rebuild = flag != 0;
    }
    if (!rebuild) return TTC_OK;
    int r = ttc_rebuild_full_at_anchors(c);
    if (r) return r;
    c->list_pruned = false;
    c->list_half = false;
    return TTC_OK;
}

// The pruned list of an MD loop in the format the loop's kernels use (half
// or full), built at the anchors in the current sorted order when the
// current list has the other format.
static int ensure_format(ttc_ctx *c, bool half) {
    if (c->have_list && c->list_half == half) return TTC_OK;
    for (int attempt = 0; attempt < 4; attempt++) {
        if (!c->have_list) {
            int r = ttc_nblist(c);
            if (r) return r;
        }
        int r = clear_overflow(c);
        if (r) return r;
        ttc_launch_build_md_at_anchors(c, half);
        TTC_HIP(hipGetLastError());
        r = read_status(c);
        if (r) return r;
        if (!(c->h_status[ST_OVERFLOW] & 11)) {
            c->list_half = half;
            c->list_pruned = true;
            return TTC_OK;
        }
        r = grow_after_overflow(c);
        if (r) return r;
    }
    ttc_set_error("list or window capacity could not be grown");
    return TTC_ERR_LIST_OVERFLOW;
}

// half lists for the velocity-Verlet loop from 256 sites on
static bool half_enabled(const ttc_ctx *c) { return ttc_half_enabled(c); }

static int finish(ttc_ctx *c, double *ff, int64_t stride) {
    TTC_HIP(hipGetLastError());
    TTC_HIP(hipMemcpyAsync(c->h_ev, c->ev, sizeof(double) * EV_N, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    if (ff) return get_xyz(c, ff, stride, c->f4, true);
    return TTC_OK;
}

extern "C" int ttc_pair(ttc_ctx *c, double *ff, int64_t stride, double *e5, double *vir6) {
    int r = need_list(c);
    if (r) return r;
    ttc_launch_pair(c, true, 0);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 1);
    r = finish(c, ff, stride);
    if (r) return r;
    if (e5) for (int i = 0; i < 5; i++) e5[i] = c->h_ev[EV_LJ + i];
    if (vir6) for (int i = 0; i < 6; i++) vir6[i] = c->h_ev[EV_PV0 + i];
    c->f4_valid = false;
    return TTC_OK;
}

extern "C" int ttc_pme_calc(ttc_ctx *c, double *ff, int64_t stride, double *energy3, double *vir6) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_pme_calc: topology not set");
    int r0 = need_list(c);   // the spread walks the sites cell by cell
    if (r0) return r0;
    ttc_launch_spread(c);
    ttc_launch_pme(c, true, true, 0);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 2 | 4);
    int r = finish(c, ff, stride);
    if (r) return r;
    if (energy3) {
        energy3[0] = c->h_ev[EV_REC];
        energy3[1] = c->e_self;
        energy3[2] = energy3[0] - energy3[1];
    }
    if (vir6) for (int i = 0; i < 6; i++) vir6[i] = c->h_ev[EV_KV0 + i];
    c->f4_valid = false;
    return TTC_OK;
}

static void slots8(const ttc_ctx *c, double *energy) {
    const double *ev = c->h_ev;
    double enk2 = ev[EV_REC] - c->e_self;
    double enr0 = ev[EV_REAL], enr1 = ev[EV_REAL] - ev[EV_REALS], enr2 = ev[EV_INTRA];
    energy[0] = ev[EV_LJ] + enr0 - enr2 + enk2;
    energy[1] = energy[0] - ev[EV_LJS];
    energy[2] = enr0 + enk2 - enr2;
    energy[3] = enr0 - enr1;
    energy[4] = ev[EV_REC];
    energy[5] = c->e_self;
    energy[6] = enr2;
    energy[7] = enr0;
}

extern "C" int ttc_cnonbondnl(ttc_ctx *c, double *ff, int64_t stride, double *energy8, double *vir_nbond6,
                              double *vir_kspace6) {
    int r = need_list(c);
    if (r) return r;
    ttc_launch_pair(c, true, 0);
    ttc_launch_spread(c);
    ttc_launch_pme(c, true, true, 0);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 1 | 2 | 4);
    r = finish(c, ff, stride);
    if (r) return r;
    if (energy8) slots8(c, energy8);
    if (vir_nbond6) for (int i = 0; i < 6; i++) vir_nbond6[i] = c->h_ev[EV_PV0 + i];
    if (vir_kspace6) for (int i = 0; i < 6; i++) vir_kspace6[i] = c->h_ev[EV_KV0 + i];
    c->f4_valid = false;
    return TTC_OK;
}

extern "C" int ttc_cbond(ttc_ctx *c, double *ff, int64_t stride, double *ebond, double *vir6) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_cbond: topology not set");
    ttc_launch_pme(c, true, false, TTC_BT_BOND);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 2);
    int r = finish(c, ff, stride);
    if (r) return r;
    if (ebond) *ebond = c->h_ev[EV_BOND];
    if (vir6) for (int i = 0; i < 6; i++) vir6[i] = c->h_ev[EV_BV0 + i];
    c->f4_valid = false;
    return TTC_OK;
}

// ------------------------------------------------- angle-type bonded terms

extern "C" int ttc_set_angles(ttc_ctx *c, int nbend, const int *bend_abc, const double *bend_eq,
                              const double *bend_k, int nurey, const int *urey_ab, const double *urey_s0,
                              const double *urey_k, int ntors, const int *tors_abcd, const double *kphi,
                              const double *nphi, const double *delphi, int nimpr, const int *impr_abcd,
                              const double *impr_k, const double *impr_eq) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_set_angles: topology not set");
    CHECK(nbend >= 0 && nurey >= 0 && ntors >= 0 && nimpr >= 0, TTC_ERR_ARG, "ttc_set_angles: negative count");
    CHECK((!nbend || (bend_abc && bend_eq && bend_k)) && (!nurey || (urey_ab && urey_s0 && urey_k)) &&
              (!ntors || (tors_abcd && kphi && nphi && delphi)) && (!nimpr || (impr_abcd && impr_k && impr_eq)),
          TTC_ERR_ARG, "ttc_set_angles: null array");
    const int n = c->n;
    const int cnt[4] = {nbend, nurey, ntors, nimpr}, ends[4] = {3, 2, 4, 4};
    const int *sites[4] = {bend_abc, urey_ab, tors_abcd, impr_abcd};
    int start[5], slot[5];
    start[0] = 0;
    slot[0] = 0;
    for (int k = 0; k < 4; k++) {
        int pad = (cnt[k] + 63) / 64 * 64;   // whole waves per kind
        start[k + 1] = start[k] + pad;
        slot[k + 1] = slot[k] + ends[k] * pad;
    }
    CHECK(slot[4] < (1 << 30), TTC_ERR_ARG, "ttc_set_angles: too many terms");
    const int ntot = start[4];
    std::vector<int4> idx(ntot, make_int4(-1, -1, -1, -1));
    std::vector<double2> par(ntot, make_double2(0.0, 0.0));
    std::vector<double> tpar((size_t)ntors * 18);
    std::vector<std::vector<int>> per(n);   // slot | kind << 30 per site, term order
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < cnt[k]; i++) {
            int s[4] = {-1, -1, -1, -1};
            for (int e = 0; e < ends[k]; e++) {
                s[e] = sites[k][(size_t)ends[k] * i + e];
                CHECK(s[e] >= 0 && s[e] < n, TTC_ERR_ARG, "ttc_set_angles: site index out of range");
            }
            for (int e = 0; e < ends[k]; e++)
                for (int f = e + 1; f < ends[k]; f++)
                    CHECK(s[e] != s[f], TTC_ERR_ARG, "ttc_set_angles: a term names one site twice");
            idx[start[k] + i] = make_int4(s[0], s[1], s[2], s[3]);
            if (k == 0) par[start[k] + i] = make_double2(bend_eq[i], bend_k[i]);
            if (k == 1) par[start[k] + i] = make_double2(urey_s0[i], urey_k[i]);
            if (k == 3) par[start[k] + i] = make_double2(impr_eq[i], impr_k[i]);
            if (k == 2)
                for (int m = 0; m < 6; m++) {
                    tpar[(size_t)i * 18 + m] = kphi[(size_t)6 * i + m];
                    tpar[(size_t)i * 18 + 6 + m] = nphi[(size_t)6 * i + m];
                    tpar[(size_t)i * 18 + 12 + m] = delphi[(size_t)6 * i + m];
                }
            for (int e = 0; e < ends[k]; e++) per[s[e]].push_back((slot[k] + ends[k] * i + e) | (k << 30));
        }
    std::vector<int> off(n + 1, 0), ref;
    for (int a = 0; a < n; a++) {
        off[a + 1] = off[a] + (int)per[a].size();
        ref.insert(ref.end(), per[a].begin(), per[a].end());
    }
    void *old[] = {c->ang_idx, c->ang_par, c->ang_tpar, c->ang_tf, c->ang_off, c->ang_ref};
    TTC_HIP(hipStreamSynchronize(c->stream));
    for (void *b : old)
        if (b) hipFree(b);
    c->ang_idx = nullptr; c->ang_par = nullptr; c->ang_tpar = nullptr;
    c->ang_tf = nullptr; c->ang_off = nullptr; c->ang_ref = nullptr;
    for (int k = 0; k < 5; k++) c->ang_start[k] = 0, c->ang_slot[k] = 0;
    c->nbend = c->nurey = c->ntors = c->nimpr = 0;
    destroy_graphs(c);
    c->f4_valid = false;
    c->mts_valid = false;
    if (ntot == 0) return TTC_OK;
    // the reductions need one partial per angle block
    int blocks = (ntot + 255) / 256;
    if (blocks > c->nblk_max) {
        hipFree(c->partials);
        c->partials = nullptr;
        c->nblk_max = blocks;
        ALLOC(c->partials, (size_t)EV_N * c->nblk_max);
        TTC_HIP(hipMemset(c->partials, 0, sizeof(double) * EV_N * c->nblk_max));
        c->gen++;
    }
    ALLOC(c->ang_idx, ntot);
    ALLOC(c->ang_par, ntot);
    ALLOC(c->ang_tpar, std::max<size_t>(tpar.size(), 1));
    ALLOC(c->ang_tf, slot[4]);
    ALLOC(c->ang_off, n + 1);
    ALLOC(c->ang_ref, std::max<size_t>(ref.size(), 1));
    TTC_HIP(hipMemcpy(c->ang_idx, idx.data(), sizeof(int4) * ntot, hipMemcpyHostToDevice));
    TTC_HIP(hipMemcpy(c->ang_par, par.data(), sizeof(double2) * ntot, hipMemcpyHostToDevice));
    if (!tpar.empty()) TTC_HIP(hipMemcpy(c->ang_tpar, tpar.data(), sizeof(double) * tpar.size(), hipMemcpyHostToDevice));
    TTC_HIP(hipMemset(c->ang_tf, 0, sizeof(double4) * slot[4]));
    TTC_HIP(hipMemcpy(c->ang_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (!ref.empty()) TTC_HIP(hipMemcpy(c->ang_ref, ref.data(), sizeof(int) * ref.size(), hipMemcpyHostToDevice));
    for (int k = 0; k < 5; k++) c->ang_start[k] = start[k], c->ang_slot[k] = slot[k];
    c->nbend = nbend; c->nurey = nurey; c->ntors = ntors; c->nimpr = nimpr;
    return TTC_OK;
}

// forces, energies (bend, urey, tors, impr), virial and (hes3 != null)
// Hessian diagonal of the angle-type kinds in bterms
static int angle_call(ttc_ctx *c, int bterms, double *ff, int64_t stride, double *e4, double *vir6, double *hes3) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc angle terms: topology not set");
    if (hes3) {
        int r = ttc_pme_theta(c);   // the Hessian gather variant stages theta even without PME
        if (r) return r;
        ttc_launch_pme_hes(c, false, bterms);
    } else {
        ttc_launch_pme(c, true, false, bterms);
    }
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 2);
    int r = finish(c, ff, stride);
    if (r) return r;
    bool any = c->ang_start[4] != 0;
    for (int k = 0; k < 4; k++)
        if (e4 && ((bterms >> (k + 1)) & 1)) e4[k] = any ? c->h_ev[EV_EBEND + k] : 0.0;
    if (vir6) for (int i = 0; i < 6; i++) vir6[i] = any ? c->h_ev[EV_AV0 + i] : 0.0;
    if (hes3) for (int d = 0; d < 3; d++) hes3[d] = any ? c->h_ev[EV_AHX + d] : 0.0;
    c->f4_valid = false;
    return TTC_OK;
}

static int angle_kind(ttc_ctx *c, int bt, double *ff, int64_t stride, double *e, double *vir6, double *hes3) {
    double e4[4] = {0.0, 0.0, 0.0, 0.0};
    int r = angle_call(c, bt, ff, stride, e4, vir6, hes3);
    if (r) return r;
    int k = bt == TTC_BT_BEND ? 0 : bt == TTC_BT_UREY ? 1 : bt == TTC_BT_TORS ? 2 : 3;
    if (e) *e = e4[k];
    return TTC_OK;
}

extern "C" int ttc_cbend(ttc_ctx *c, double *ff, int64_t stride, double *ebend, double *vir6) {
    return angle_kind(c, TTC_BT_BEND, ff, stride, ebend, vir6, nullptr);
}
extern "C" int ttc_curey(ttc_ctx *c, double *ff, int64_t stride, double *eurey, double *vir6) {
    return angle_kind(c, TTC_BT_UREY, ff, stride, eurey, vir6, nullptr);
}
extern "C" int ttc_ctorsion(ttc_ctx *c, double *ff, int64_t stride, double *etors, double *vir6) {
    return angle_kind(c, TTC_BT_TORS, ff, stride, etors, vir6, nullptr);
}
extern "C" int ttc_cimproper(ttc_ctx *c, double *ff, int64_t stride, double *eimpr, double *vir6) {
    return angle_kind(c, TTC_BT_IMPR, ff, stride, eimpr, vir6, nullptr);
}
extern "C" int ttc_cbend2(ttc_ctx *c, double *ff, int64_t stride, double *ebend, double *hes3) {
    CHECK(hes3, TTC_ERR_ARG, "ttc_cbend2: null hes3");
    return angle_kind(c, TTC_BT_BEND, ff, stride, ebend, nullptr, hes3);
}
extern "C" int ttc_curey2(ttc_ctx *c, double *ff, int64_t stride, double *eurey, double *hes3) {
    CHECK(hes3, TTC_ERR_ARG, "ttc_curey2: null hes3");
    return angle_kind(c, TTC_BT_UREY, ff, stride, eurey, nullptr, hes3);
}
extern "C" int ttc_ctorsion2(ttc_ctx *c, double *ff, int64_t stride, double *etors, double *hes3) {
    CHECK(hes3, TTC_ERR_ARG, "ttc_ctorsion2: null hes3");
    return angle_kind(c, TTC_BT_TORS, ff, stride, etors, nullptr, hes3);
}
extern "C" int ttc_cimproper2(ttc_ctx *c, double *ff, int64_t stride, double *eimpr, double *hes3) {
    CHECK(hes3, TTC_ERR_ARG, "ttc_cimproper2: null hes3");
    return angle_kind(c, TTC_BT_IMPR, ff, stride, eimpr, nullptr, hes3);
}

extern "C" int ttc_bonded_energies(ttc_ctx *c, double *e5) {
    CHECK(c && e5, TTC_ERR_ARG, "ttc_bonded_energies: bad argument");
    bool any = c->ang_start[4] != 0;
    e5[0] = c->h_ev[EV_BOND];
    for (int k = 0; k < 4; k++) e5[1 + k] = any ? c->h_ev[EV_EBEND + k] : 0.0;
    return TTC_OK;
}

// ------------------------------------------------------------- CONFIGT

extern "C" int ttc_cnonbondnl2(ttc_ctx *c, double *ff, int64_t stride, double *energy8, double *hes3) {
    int r = need_list(c);
    if (r) return r;
    r = ttc_pme_theta(c);
    if (r) return r;
    ttc_launch_pair_hes(c);
    ttc_launch_spread(c);
    ttc_launch_pme_hes(c, true, 0);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 1 | 2 | 4);
    r = finish(c, ff, stride);
    if (r) return r;
    if (energy8) slots8(c, energy8);
    if (hes3) for (int d = 0; d < 3; d++) hes3[d] = c->h_ev[EV_PHX + d] + c->h_ev[EV_KHX + d];
    c->f4_valid = false;
    return TTC_OK;
}

extern "C" int ttc_cbond2(ttc_ctx *c, double *ff, int64_t stride, double *ebond, double *hes3) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_cbond2: topology not set");
    int r = ttc_pme_theta(c);   // the Hessian gather variant stages theta even without PME
    if (r) return r;
    ttc_launch_pme_hes(c, false, TTC_BT_BOND);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 2);
    r = finish(c, ff, stride);
    if (r) return r;
    if (ebond) *ebond = c->h_ev[EV_BOND];
    if (hes3) for (int d = 0; d < 3; d++) hes3[d] = c->h_ev[EV_BHX + d];
    c->f4_valid = false;
    return TTC_OK;
}

extern "C" int ttc_configtemp(ttc_ctx *c, double *num, double *den, double *T, double *hes3) {
    int r = need_list(c);
    if (r) return r;
    r = ttc_pme_theta(c);
    if (r) return r;
    ttc_launch_pair_hes(c);
    ttc_launch_spread(c);
    ttc_launch_pme_hes(c, true, TTC_BT_ALL);
    ttc_launch_reduce(c, false);
    ttc_launch_combine(c, 7);
    TTC_HIP(hipGetLastError());
    std::vector<double4> f(c->n);
    TTC_HIP(hipMemcpyAsync(c->h_ev, c->ev, sizeof(double) * EV_N, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipMemcpyAsync(f.data(), c->f4, sizeof(double4) * c->n, hipMemcpyDeviceToHost, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    double f2 = 0.0;
    for (const double4 &v : f) f2 += v.x * v.x + v.y * v.y + v.z * v.z;
    double h[3];
    for (int d = 0; d < 3; d++)
        h[d] = c->h_ev[EV_PHX + d] + c->h_ev[EV_KHX + d] + c->h_ev[EV_BHX + d] + (c->ang_start[4] ? c->h_ev[EV_AHX + d] : 0.0);
    double hr = h[0] + h[1] + h[2];
    if (num) *num = f2;
    if (den) *den = hr;
    if (T) *T = f2 / hr / (TTC_KB * TTC_NA) / 0.001;   // configtemp.c:43-45
    if (hes3) for (int d = 0; d < 3; d++) hes3[d] = h[d];
    c->f4_valid = true;   // f4 = the full force at the current positions
    return TTC_OK;
}

// one full force evaluation into f4 (+ energies in ev)
static void launch_forces(ttc_ctx *c, bool energy) {
    ttc_launch_pair(c, energy, 0);
    ttc_launch_spread(c);
    ttc_launch_pme(c, energy, true, TTC_BT_ALL);
    ttc_launch_combine(c, 7);
    if (energy) ttc_launch_reduce(c, false);
}

extern "C" int ttc_forces(ttc_ctx *c, double *ff, int64_t stride, double *en9) {
    int r = need_list(c);
    if (r) return r;
    launch_forces(c, true);
    if (ff) {
        // forces.c:126-130 zeroes ff first
        for (int i = 0; i < c->n; i++) { double *d = ff + (size_t)i * stride; d[0] = d[1] = d[2] = 0.0; }
    }
    r = finish(c, ff, stride);
    if (r) return r;
    c->f4_valid = true;
    if (en9) {
        double e8[8];
        slots8(c, e8);
        en9[0] = c->h_ev[EV_BOND];
        en9[1] = e8[0]; en9[2] = e8[1]; en9[3] = e8[2]; en9[4] = e8[7];
        en9[5] = e8[3]; en9[6] = e8[4]; en9[7] = e8[5]; en9[8] = e8[6];
    }
    return TTC_OK;
}

extern "C" int ttc_drift(ttc_ctx *c) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_drift: topology not set");
    ttc_launch_momentum_removal(c);
    TTC_HIP(hipGetLastError());
    TTC_HIP(hipStreamSynchronize(c->stream));
    return TTC_OK;
}

extern "C" int ttc_set_drift(ttc_ctx *c, int every) {
    CHECK(c && every >= 0, TTC_ERR_ARG, "ttc_set_drift: bad argument");
    c->drift_every = every;
    return TTC_OK;
}

extern "C" int ttc_kinet(ttc_ctx *c, double *kinet, double *ken6) {
    CHECK(c, TTC_ERR_ARG, "ttc_kinet: null context");
    ttc_launch_kinet(c);
    ttc_launch_reduce(c, false);
    int r = finish(c, nullptr, 0);
    if (r) return r;
    if (kinet) *kinet = c->h_ev[EV_KE0] + c->h_ev[EV_KE2] + c->h_ev[EV_KE5];
    if (ken6) for (int i = 0; i < 6; i++) ken6[i] = c->h_ev[EV_KE0 + i];
    return TTC_OK;
}

// One velocity-Verlet step as launched (and captured): kick + drift,
// conditional rebuild, pair, spread + PME (+ bonds), kick + trigger.
// integrate_vverlet.c:36-103 with the DLIST trigger of nve_md.c:369-370:
// the drift checks the displacement and the conditional rebuild runs
// before this step's forces, with the same anchors (this step's positions)
// the reference rebuilds at after the step.
// The spread reads copies of the drifted positions and of the cell ranges
// the drift writes, the FFTs only the charge grid, and the gather works in
// the original numbering on a snapshot the drift writes, so all three run
// on the side stream while the main stream rebuilds the list and runs the
// pair kernel: one fork and one join per step.
void ttc_side_fork(ttc_ctx *c, int k) {
    hipEventRecord(c->sev[k], c->stream);
    hipStreamWaitEvent(c->side, c->sev[k], 0);
}

void ttc_side_join(ttc_ctx *c) {
    hipEventRecord(c->sev[2], c->side);
    hipStreamWaitEvent(c->stream, c->sev[2], 0);
}

//
// The second half kick of a plain step is deferred into the next step's
// drift kernel (head = true), which saves a dependent launch per step; an
// energy step ends with its own kick (the kinetic energy of the row) and the
// run with one more kick, so every call leaves the state a full step ends
// with.  The pair kernel counts the step and clears the trigger flag.
static void launch_step(ttc_ctx *c, double vscal, bool energy, bool head) {
    ttc_launch_md_drift(c, vscal, head);
    // the main stream's rebuild and pair kernel are issued before the side
    // stream's work, so the longer chain is launched first
    hipEventRecord(c->sev[0], c->stream);
    ttc_launch_rebuild(c, true);
    ttc_launch_pair_step(c, energy);
    hipStreamWaitEvent(c->side, c->sev[0], 0);
    {
        TtcOnSide side(c);
        ttc_launch_md_spread(c);
        ttc_launch_fft(c, energy);
        ttc_launch_gather_orig(c, energy);
    }
    ttc_side_join(c);
    if (energy) {
        ttc_launch_kick_tail(c, vscal, true);
        ttc_launch_reduce(c, true);
    }
}

static int capture(ttc_ctx *c, double vscal, bool energy, bool head, hipGraphExec_t *exec, int nsteps = 1) {
    hipGraph_t g;
    TTC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < nsteps; k++) launch_step(c, vscal, energy, head);
    TTC_HIP(hipStreamEndCapture(c->stream, &g));
    TTC_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    return TTC_OK;
}

// The tail of a step whose conditional rebuild overflowed: the kernels
// after that rebuild stopped (the step's drift is done, positions x_k), so
// the list is grown and built at x_k -- the build the step made -- and the
// rest of the step runs: spreading x_k again (the step's FFT consumed the
// grid), pair kernel, FFTs, gather, second half kick and the energy row.
static int redo_overflowed_step(ttc_ctx *c, double vscal, bool energy) {
    for (int attempt = 0; attempt < 4; attempt++) {
        int r = grow_after_overflow(c);
        if (r) return r;
        r = clear_overflow(c);
        if (r) return r;
        ttc_launch_rebuild_md(c);
        TTC_HIP(hipGetLastError());
        r = read_status(c);
        if (r) return r;
        if (!(c->h_status[ST_OVERFLOW] & 11)) {
            ttc_launch_spread(c);
            ttc_launch_pair_step(c, energy);
            ttc_launch_fft(c, energy);
            ttc_launch_gather_orig(c, energy);
            ttc_launch_kick_tail(c, vscal, energy);
            if (energy) ttc_launch_reduce(c, true);
            TTC_HIP(hipGetLastError());
            return TTC_OK;
        }
    }
    ttc_set_error("ttc_md_vverlet: list capacity could not be grown");
    return TTC_ERR_LIST_OVERFLOW;
}

static int capture_vv_graphs(ttc_ctx *c, double vscal) {
    destroy_graphs(c);
    for (int h = 0; h < 2; h++)
        for (int e = 0; e < 2; e++) {
            int r = capture(c, vscal, e == 1, h == 1, &c->g_vv[h][e]);
            if (r) return r;
        }
    int r = capture(c, vscal, false, true, &c->g_vv_run, TTC_VV_RUN);
    if (r) return r;
    c->g_vscal = vscal;
    return TTC_OK;
}

extern "C" int ttc_md_vverlet(ttc_ctx *c, int nsteps, double vscal, int energy_every, double *energies,
                              int max_rows, int *nrows, int *rebuilds) {
    CHECK(c && c->topo && nsteps >= 0, TTC_ERR_ARG, "ttc_md_vverlet: bad argument");
    CHECK(energy_every == 0 || (energies && max_rows > 0), TTC_ERR_ARG, "ttc_md_vverlet: no row buffer");
    const bool half = half_enabled(c);
    if (c->half_mode != half) {
        c->half_mode = half;
        destroy_graphs(c);
    }
    int r = need_list(c, half);
    if (r) return r;
    if (energy_every > 0 && max_rows > c->rows_cap) {
        if (c->rows) hipFree(c->rows);
        ALLOC(c->rows, (size_t)max_rows * ROW_LEN);
        c->rows_cap = max_rows;
        destroy_graphs(c);
    }
    if (!c->f4_valid) {
        // main.c:489-493 / nve_md.c:301-303: forces before the first step
        launch_forces(c, false);
        c->f4_valid = true;
    }
    r = ensure_format(c, half);
    if (r) return r;
    // counters; the trigger for step 0 is checked on the current positions
    TTC_HIP(hipMemsetAsync(c->status, 0, sizeof(int) * ST_NSTAT, c->stream));
    ttc_launch_nlcheck(c);
    EventPair ev;
    TTC_HIP(ev.create());
    TTC_HIP(hipEventRecord(ev.e0, c->stream));
    bool kicked = true;   // the state ends a full step: no kick pending
    int step = 1;
    for (int redo = 0;; redo++) {
        if (!c->g_vv[0][0] || c->g_vscal != vscal) {
            r = capture_vv_graphs(c, vscal);
            if (r) return r;
        }
        while (step <= nsteps) {
            // TTC_VV_RUN plain steps in one graph (no launch gap between
            // them) when a kick is pending and none of them writes an energy row
            int last = step + TTC_VV_RUN - 1;
            const int de = c->drift_every;
            bool run = !kicked && last <= nsteps &&
                       (energy_every <= 0 || last / energy_every == (step - 1) / energy_every) &&
                       (de <= 0 || last / de == (step - 1) / de);
            if (run) {
                TTC_HIP(hipGraphLaunch(c->g_vv_run, c->stream));
                step = last + 1;
                continue;
            }
            bool en = energy_every > 0 && step % energy_every == 0;
            TTC_HIP(hipGraphLaunch(c->g_vv[kicked ? 0 : 1][en ? 1 : 0], c->stream));
            kicked = en;
            if (de > 0 && step % de == 0) {
                // nve_md.c:208: drift(k) after the step's kinet / calcvalue
                if (!kicked) ttc_launch_kick_tail(c, vscal, false);
                ttc_launch_momentum_removal(c);
                kicked = true;
            }
            step++;
        }
        if (!kicked) ttc_launch_kick_tail(c, vscal, false);
        kicked = true;
        r = read_status(c);
        if (r) return r;
        if (!(c->h_status[ST_OVERFLOW] & TTC_OVF_REDO)) break;
        // a rebuild overflowed: ST_STEP steps completed, the next one
        // stopped after its drift; its remaining kernels run here, then the
        // loop continues after it (graphs are re-captured if buffers grew)
        CHECK(redo < 16, TTC_ERR_LIST_OVERFLOW, "ttc_md_vverlet: repeated list overflow");
        const int failed = c->h_status[ST_STEP] + 1;
        r = redo_overflowed_step(c, vscal, energy_every > 0 && failed % energy_every == 0);
        if (r) return r;
        if (c->drift_every > 0 && failed % c->drift_every == 0) ttc_launch_momentum_removal(c);
        step = failed + 1;
    }
    TTC_HIP(hipEventRecord(ev.e1, c->stream));
    r = read_status(c);
    if (r) return r;
    float ms = 0.0f;
    TTC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
    c->t_step = nsteps > 0 ? 1000.0 * ms / nsteps : 0.0;
    int nr = std::min(c->h_status[ST_NROWS], max_rows);
    if (nrows) *nrows = energy_every > 0 ? nr : 0;
    if (rebuilds) *rebuilds = c->h_status[ST_NREBUILD];
    if (c->h_status[ST_NREBUILD] > 0) c->list_pruned = true;
    if (energy_every > 0 && nr > 0)
        TTC_HIP(hipMemcpy(energies, c->rows, sizeof(double) * ROW_LEN * nr, hipMemcpyDeviceToHost));
    c->mts_valid = false;
    if (c->h_status[ST_OVERFLOW] & 16) {
        ttc_set_error("ttc_md_vverlet: a pair force beyond 6.7e7 kJ/mol/A (overlapping sites)");
        return TTC_ERR_ARG;
    }
    if (c->h_status[ST_OVERFLOW]) {
        ttc_set_error("ttc_md_vverlet: neighbour list overflow during the run (nblist.c:185-196)");
        return TTC_ERR_LIST_OVERFLOW;
    }
    return TTC_OK;
}

extern "C" int ttc_timing(ttc_ctx *c, int which, double *usec) {
    CHECK(c && usec, TTC_ERR_ARG, "ttc_timing: bad argument");
    if (which == 2) { *usec = c->t_step; return TTC_OK; }
    CHECK(which >= 0 && which <= 6, TTC_ERR_ARG, "ttc_timing: which must be 0 .. 6");
    int r = need_list(c, true);   // the pair kernel of the current list format
    if (r) return r;
    if (which >= 5) {
        // the MD loop's list build at the anchors, half (5) or full (6)
        // format; afterwards the list is rebuilt in its own format
        const bool had_half = c->list_half;
        c->list_half = !(which == 5);   // ensure_format builds when the format differs
        r = ensure_format(c, which == 5);
        if (r) return r;
        EventPair ev;
        TTC_HIP(ev.create());
        TTC_HIP(hipEventRecord(ev.e0, c->stream));
        for (int i = 0; i < 10; i++) ttc_launch_build_md_at_anchors(c, which == 5);
        TTC_HIP(hipEventRecord(ev.e1, c->stream));
        TTC_HIP(hipEventSynchronize(ev.e1));
        float ms = 0.0f;
        TTC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
        *usec = 1000.0 * ms / 10;
        c->list_half = which == 5;
        return ensure_format(c, had_half);
    }
    const int reps = 20;
    EventPair ev;
    TTC_HIP(ev.create());
    // the gather in the original numbering reads the positions snapshot an
    // MD step's drift writes: take it from the current positions
    if (which == 4) ttc_launch_snap(c);
    // warm-up launch, then timed back-to-back launches on the context stream
    // (the pair kernel alone: its fixed-point coordinates are made once)
    if (which == 0) ttc_launch_posq(c);
    auto launch = [&] {
        if (which == 0) ttc_launch_pair_only(c);
        else if (which == 1) ttc_launch_spread(c);
        else if (which == 3) ttc_launch_fft(c, false);   // the five line passes
        else ttc_launch_gather_orig(c, false);
    };
    launch();
    TTC_HIP(hipEventRecord(ev.e0, c->stream));
    for (int i = 0; i < reps; i++) launch();
    TTC_HIP(hipEventRecord(ev.e1, c->stream));
    TTC_HIP(hipEventSynchronize(ev.e1));
    float ms = 0.0f;
    TTC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
    if (which == 0) {
        // the repeated half-list launches accumulated partner forces
        TTC_HIP(hipMemsetAsync(c->fixj, 0, sizeof(long long) * 3 * c->n, c->stream));
        TTC_HIP(hipStreamSynchronize(c->stream));
    }
    if (which == 1) {
        // the repeated spreads accumulated: clear the grid again
        TTC_HIP(hipMemsetAsync(c->qfix, 0, sizeof(long long) * c->ngrid, c->stream));
        TTC_HIP(hipStreamSynchronize(c->stream));
    }
    // the FFT passes leave the grid cleared and the gather overwrote the PME
    // forces with those of that grid: forces are recomputed before the next
    // MD call
    if (which >= 3) c->f4_valid = false;
    *usec = 1000.0 * ms / reps;
    if (which == 0) c->t_pair = *usec; else if (which == 1) c->t_spread = *usec;
    return TTC_OK;
}

// Times the pair kernel at 4, 8 and 16 waves per 64-site group on the
// current list (ttc_timing(0): HIP events around 20 launches) and keeps the
// fastest; captured graphs are re-made if the choice changes.
int ttc_spread_chunk_ok(const ttc_ctx *c, int chunk);
int ttc_spread_attrs(ttc_ctx *c);

// the spread chunk: 64, 128 and 256 sites per pass where they fit, timed
// with ttc_timing(1); the fastest stays
static int tune_spread(ttc_ctx *c, double *usec3) {
    const int cand[3] = {64, 128, 256};
    const int old = c->spread_chunk;
    int best = old;
    double tbest = 1e300;
    for (int i = 0; i < 3; i++) {
        if (usec3) usec3[i] = 0.0;
        if (!ttc_spread_chunk_ok(c, cand[i])) continue;
        c->spread_chunk = cand[i];
        int r = ttc_spread_attrs(c);
        double t = 0.0;
        if (!r) r = ttc_timing(c, 1, &t);
        if (r) { c->spread_chunk = old; ttc_spread_attrs(c); return r; }
        if (usec3) usec3[i] = t;
        if (t < tbest) { tbest = t; best = cand[i]; }
    }
    c->spread_chunk = best;
    c->t_spread = tbest;
    return ttc_spread_attrs(c);
}

extern "C" int ttc_tune(ttc_ctx *c, int *pair_waves, double *usec3) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_tune: topology not set");
    const int cand[3] = {4, 8, 16};
    // the pair kernel of the current list (the half-list kernel after a
    // velocity-Verlet loop, the full-list kernel otherwise)
    int r0 = need_list(c, true);
    if (r0) return r0;
    int &waves = c->list_half ? c->pair_waves_h : c->pair_waves;
    const int ncand = c->list_half ? 2 : 3;   // the half-list kernel runs 4 or 8 waves
    int old = waves, best = old;
    double tbest = 1e300;
    if (usec3) usec3[2] = 0.0;
    for (int i = 0; i < ncand; i++) {
        waves = cand[i];
        double t = 0.0;
        int r = ttc_timing(c, 0, &t);
        if (r) { waves = old; return r; }
        if (usec3) usec3[i] = t;
        if (t < tbest) { tbest = t; best = cand[i]; }
    }
    waves = best;
    c->t_pair = tbest;
    if (best != old) {
        destroy_graphs(c);
        c->gen++;
    }
    if (pair_waves) *pair_waves = best;
    return TTC_OK;
}

extern "C" int ttc_tune_spread(ttc_ctx *c, int *chunk, double *usec3) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_tune_spread: topology not set");
    const int old = c->spread_chunk;
    int r = tune_spread(c, usec3);
    if (r) return r;
    if (c->spread_chunk != old) {
        destroy_graphs(c);
        c->gen++;
    }
    if (chunk) *chunk = c->spread_chunk;
    return TTC_OK;
}

// A line of the tuning cache: the key of a context (sizes, box, grid,
// cutoffs, list format, type count, GPU) and the choices made for it.
static std::string tune_key(ttc_ctx *c) {
    hipDeviceProp_t prop;
    int dev = 0;
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
    }
    for (char &ch : arch)
        if (ch == ' ' || ch == '\t') ch = '_';
    char buf[512];
    snprintf(buf, sizeof buf, "n=%d box=%.9g,%.9g,%.9g grid=%d,%d,%d order=%d rc=%.9g rclist=%.9g half=%d ntypes=%d gpu=%s/%d",
             c->n, c->p.lx, c->p.ly, c->p.lz, c->p.gx, c->p.gy, c->p.gz, c->p.order, c->p.rc, c->p.rclist,
             c->list_half ? 1 : 0, c->ntypes, arch.c_str(), cus);
    return buf;
}

extern "C" int ttc_tune_cached(ttc_ctx *c, const char *path, int *pair_waves, int *spread_chunk, int *from_cache) {
    CHECK(c && c->topo && path, TTC_ERR_ARG, "ttc_tune_cached: bad argument");
    int r = need_list(c, true);
    if (r) return r;
    const std::string key = tune_key(c);
    int w = 0, ch = 0, hit = 0;
    if (FILE *f = fopen(path, "r")) {
        char line[1024];
        while (fgets(line, sizeof line, f)) {
            char *tab = strchr(line, '\t');
            if (!tab) continue;
            *tab = 0;
            int lw = 0, lc = 0;
            if (key == line && sscanf(tab + 1, "%d %d", &lw, &lc) == 2 && (lw == 4 || lw == 8 || lw == 16) &&
                (lc == 64 || lc == 128 || lc == 256) && !(c->list_half && lw == 16) && ttc_spread_chunk_ok(c, lc)) {
                w = lw; ch = lc; hit = 1;   // the last valid line of the key wins
            }
        }
        fclose(f);
    }
    if (hit) {
        int &waves = c->list_half ? c->pair_waves_h : c->pair_waves;
        const bool changed = waves != w || c->spread_chunk != ch;
        waves = w;
        c->spread_chunk = ch;
        r = ttc_spread_attrs(c);
        if (r) return r;
        if (changed) {
            destroy_graphs(c);
            c->gen++;
        }
    } else {
        r = ttc_tune(c, &w, nullptr);
        if (!r) r = ttc_tune_spread(c, &ch, nullptr);
        if (r) return r;
        if (FILE *f = fopen(path, "a")) {
            fprintf(f, "%s\t%d %d\n", key.c_str(), w, ch);
            fclose(f);
        }
    }
    if (pair_waves) *pair_waves = w;
    if (spread_chunk) *spread_chunk = ch;
    if (from_cache) *from_cache = hit;
    return TTC_OK;
}

extern "C" int ttc_list_stats(ttc_ctx *c, int64_t *entries, int *max_row, int *capacity) {
    CHECK(c, TTC_ERR_ARG, "ttc_list_stats: null context");
    TTC_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> gl(c->G), gc(c->G);
    TTC_HIP(hipMemcpy(gl.data(), c->glen, sizeof(int) * c->G, hipMemcpyDeviceToHost));
    TTC_HIP(hipMemcpy(gc.data(), c->gcnt, sizeof(int) * c->G, hipMemcpyDeviceToHost));
    int64_t tot = 0;
    int mx = 0;
    for (int g = 0; g < c->G; g++) { tot += gc[g]; mx = std::max(mx, gl[g]); }
    if (entries) *entries = tot;   // entries of the device list (each pair in both rows)
    if (max_row) *max_row = mx;
    if (capacity) *capacity = c->cap;
    return TTC_OK;
}

extern "C" int ttc_get_forces(ttc_ctx *c, double *ff, int64_t stride) {
    CHECK(c && ff && stride >= 3, TTC_ERR_ARG, "ttc_get_forces: bad argument");
    CHECK(c->f4_valid, TTC_ERR_ARG, "ttc_get_forces: no force of the current positions (run a step or ttc_forces)");
    return get_xyz(c, ff, stride, c->f4, false);
}

extern "C" int ttc_set_capacity_limits(ttc_ctx *c, int row_limit, int window_limit) {
    CHECK(c && row_limit >= 0 && window_limit >= 0, TTC_ERR_ARG, "ttc_set_capacity_limits: bad argument");
    c->row_limit = row_limit;
    c->win_limit = window_limit;
    destroy_graphs(c);
    return TTC_OK;
}

extern "C" int ttc_window_stats(ttc_ctx *c, int *max_window, int *window_capacity, int *half) {
    CHECK(c, TTC_ERR_ARG, "ttc_window_stats: null context");
    TTC_HIP(hipStreamSynchronize(c->stream));
    int wl = 0;
    if (c->list_half) {
        std::vector<int> w(c->G);
        TTC_HIP(hipMemcpy(w.data(), c->wlen, sizeof(int) * c->G, hipMemcpyDeviceToHost));
        for (int g = 0; g < c->G; g++) wl = std::max(wl, w[g]);
    }
    if (max_window) *max_window = wl;
    if (window_capacity) *window_capacity = c->wcap;
    if (half) *half = c->list_half ? 1 : 0;
    return TTC_OK;
}

// ------------------------------------------------------------------ RESPA

static void launch_force_short(ttc_ctx *c, bool energy) {
    ttc_launch_pme(c, energy, false, TTC_BT_SHORT);   // force_short.c: bonds, bends, Urey-Bradley
    ttc_launch_combine_to(c, 2, c->fs);
}

static void launch_force_long(ttc_ctx *c, bool energy) {
    ttc_launch_pair(c, energy, 0);
    ttc_launch_spread(c);
    ttc_launch_pme(c, energy, true, TTC_BT_LONG);   // force_long.c: PME, torsions, impropers
    ttc_launch_combine_to(c, 1 | 2 | 4, c->fl);
}

extern "C" int ttc_force_short(ttc_ctx *c, double *ff_short, int64_t stride, double *ebond) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_force_short: topology not set");
    launch_force_short(c, true);
    ttc_launch_reduce(c, false);
    int r = finish(c, nullptr, 0);
    if (r) return r;
    if (ff_short) {
        for (int i = 0; i < c->n; i++) { double *d = ff_short + (size_t)i * stride; d[0] = d[1] = d[2] = 0.0; }
        r = get_xyz(c, ff_short, stride, c->fs, true);
        if (r) return r;
    }
    if (ebond) *ebond = c->h_ev[EV_BOND];
    return TTC_OK;
}

extern "C" int ttc_force_long(ttc_ctx *c, double *ff_long, int64_t stride, double *energy8) {
    int r = need_list(c);
    if (r) return r;
    launch_force_long(c, true);
    ttc_launch_reduce(c, false);
    r = finish(c, nullptr, 0);
    if (r) return r;
    if (ff_long) {
        for (int i = 0; i < c->n; i++) { double *d = ff_long + (size_t)i * stride; d[0] = d[1] = d[2] = 0.0; }
        r = get_xyz(c, ff_long, stride, c->fl, true);
        if (r) return r;
    }
    if (energy8) slots8(c, energy8);
    return TTC_OK;
}

// one RESPA long step (integrate_mts.c:38-140); the DLIST trigger is
// checked after the short steps and the conditional rebuild runs before
// the long forces of the new positions
static void launch_mts_step(ttc_ctx *c, int ninner, double dtlong, double vscal, bool energy) {
    double dts = c->p.dt;
    ttc_launch_vkick(c, c->fl, vscal, vscal * 0.5 * dtlong, false, false, false);
    for (int j = 0; j < ninner; j++) {
        ttc_launch_vkick_drift(c, c->fs, 1.0, vscal * 0.5 * dts, dts);
        launch_force_short(c, energy && j == ninner - 1);
        ttc_launch_vkick(c, c->fs, 1.0, vscal * 0.5 * dts, false, false, false);
    }
    ttc_launch_nlcheck(c);
    ttc_launch_rebuild(c, true);
    launch_force_long(c, energy);
    ttc_launch_vkick(c, c->fl, 1.0, vscal * 0.5 * dtlong, false, energy, true);
    if (energy) ttc_launch_reduce(c, true);
}

static int capture_fn(ttc_ctx *c, hipGraphExec_t *exec, const std::function<void()> &body) {
    hipGraph_t g;
    TTC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    body();
    TTC_HIP(hipStreamEndCapture(c->stream, &g));
    TTC_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    return TTC_OK;
}

static int run_graphs(ttc_ctx *c, hipGraphExec_t plain, hipGraphExec_t energy, int nsteps, int energy_every,
                      double *energies, int max_rows, int *nrows, int *rebuilds) {
    TTC_HIP(hipMemsetAsync(c->status, 0, sizeof(int) * ST_NSTAT, c->stream));
    ttc_launch_nlcheck(c);
    EventPair ev;
    TTC_HIP(ev.create());
    TTC_HIP(hipEventRecord(ev.e0, c->stream));
    for (int step = 1; step <= nsteps; step++) {
        bool en = energy_every > 0 && step % energy_every == 0;
        TTC_HIP(hipGraphLaunch(en ? energy : plain, c->stream));
        // nvt_md.c:302: drift(k) after the step
        if (c->drift_every > 0 && step % c->drift_every == 0) ttc_launch_momentum_removal(c);
    }
    TTC_HIP(hipEventRecord(ev.e1, c->stream));
    int r = read_status(c);
    if (r) return r;
    float ms = 0.0f;
    TTC_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
    c->t_step = nsteps > 0 ? 1000.0 * ms / nsteps : 0.0;
    int nr = std::min(c->h_status[ST_NROWS], max_rows);
    if (nrows) *nrows = energy_every > 0 ? nr : 0;
    if (rebuilds) *rebuilds = c->h_status[ST_NREBUILD];
    if (c->h_status[ST_NREBUILD] > 0) c->list_pruned = true;
    if (energy_every > 0 && nr > 0)
        TTC_HIP(hipMemcpy(energies, c->rows, sizeof(double) * ROW_LEN * nr, hipMemcpyDeviceToHost));
    if (c->h_status[ST_OVERFLOW]) {
        ttc_set_error("neighbour list overflow during the run (nblist.c:185-196)");
        return TTC_ERR_LIST_OVERFLOW;
    }
    return TTC_OK;
}

static int ensure_rows(ttc_ctx *c, int energy_every, int max_rows) {
    if (energy_every > 0 && max_rows > c->rows_cap) {
        if (c->rows) hipFree(c->rows);
        ALLOC(c->rows, (size_t)max_rows * ROW_LEN);
        c->rows_cap = max_rows;
        destroy_graphs(c);
        c->gen++;
    }
    return TTC_OK;
}

// the RESPA loop keeps full lists
static void full_mode(ttc_ctx *c) {
    if (c->half_mode) {
        c->half_mode = false;
        destroy_graphs(c);
    }
}

extern "C" int ttc_md_mts(ttc_ctx *c, int nouter, int ninner, double dtlong, double vscal, int energy_every,
                          double *energies, int max_rows, int *nrows, int *rebuilds) {
    CHECK(c && c->topo && nouter >= 0 && ninner >= 1 && dtlong > 0, TTC_ERR_ARG, "ttc_md_mts: bad argument");
    CHECK(energy_every == 0 || (energies && max_rows > 0), TTC_ERR_ARG, "ttc_md_mts: no row buffer");
    full_mode(c);
    int r = need_list(c);
    if (r) return r;
    r = ensure_rows(c, energy_every, max_rows);
    if (r) return r;
    if (!c->mts_valid) {
        // main.c:489: force_short; force_long before the loop
        launch_force_short(c, false);
        launch_force_long(c, false);
        c->mts_valid = true;
    }
    if (!c->g_mts_plain || c->g_mts_ninner != ninner || c->g_mts_dtlong != dtlong || c->g_mts_vscal != vscal) {
        for (hipGraphExec_t *g : {&c->g_mts_plain, &c->g_mts_energy})
            if (*g) { hipGraphExecDestroy(*g); *g = nullptr; }
        r = capture_fn(c, &c->g_mts_plain, [&] { launch_mts_step(c, ninner, dtlong, vscal, false); });
        if (r) return r;
        r = capture_fn(c, &c->g_mts_energy, [&] { launch_mts_step(c, ninner, dtlong, vscal, true); });
        if (r) return r;
        c->g_mts_ninner = ninner; c->g_mts_dtlong = dtlong; c->g_mts_vscal = vscal;
    }
    r = run_graphs(c, c->g_mts_plain, c->g_mts_energy, nouter, energy_every, energies, max_rows, nrows, rebuilds);
    c->f4_valid = false;
    return r;
}

// ------------------------------------------------------------- Nose-Hoover

extern "C" int ttc_set_nhc(ttc_ctx *c, double T, double Q, int M, int Nc, int Nys, const double *w, double nfree) {
    CHECK(c && T > 0 && Q > 0 && M >= 1 && M <= 8 && Nc >= 1 && Nys >= 1 && Nys <= 8 && nfree > 0, TTC_ERR_ARG,
          "ttc_set_nhc: bad argument");
    c->nhc_T = T; c->nhc_Q = Q; c->nhc_M = M; c->nhc_Nc = Nc; c->nhc_Nys = Nys; c->nhc_nfree = nfree;
    for (int i = 0; i < 8; i++) c->nhc_w[i] = i < Nys ? (w ? w[i] : (Nys == 1 ? 1.0 : 0.0)) : 0.0;
    double st[TTC_NHC_LEN] = {0};
    st[TTC_NHC_SCALE] = 1.0;
    st[TTC_NHC_T] = T;
    st[TTC_NHC_Q] = Q;
    TTC_HIP(hipMemcpyAsync(c->nhc, st, sizeof st, hipMemcpyHostToDevice, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    for (hipGraphExec_t *g : {&c->g_nhc_plain, &c->g_nhc_energy})
        if (*g) { hipGraphExecDestroy(*g); *g = nullptr; }
    c->gen++;
    return TTC_OK;
}

extern "C" int ttc_set_nhc_temperature(ttc_ctx *c, double T, double Q) {
    CHECK(c && c->nhc_M > 0 && T > 0 && Q > 0, TTC_ERR_ARG, "ttc_set_nhc_temperature: thermostat not set");
    double tq[2] = {T, Q};
    TTC_HIP(hipMemcpyAsync(c->nhc + TTC_NHC_T, tq, sizeof tq, hipMemcpyHostToDevice, c->stream));
    TTC_HIP(hipStreamSynchronize(c->stream));
    c->nhc_T = T;
    c->nhc_Q = Q;
    return TTC_OK;
}

extern "C" int ttc_scale_velocities(ttc_ctx *c, double factor) {
    CHECK(c && c->topo, TTC_ERR_ARG, "ttc_scale_velocities: topology not set");
    TTC_HIP(hipMemcpyAsync(c->nhc + TTC_NHC_VMUL, &factor, sizeof factor, hipMemcpyHostToDevice, c->stream));
    ttc_launch_vmul(c);
    TTC_HIP(hipGetLastError());
    TTC_HIP(hipStreamSynchronize(c->stream));
    return TTC_OK;
}

extern "C" int ttc_get_nhc(ttc_ctx *c, double *zeta) {
    CHECK(c && zeta && c->nhc_M > 0, TTC_ERR_ARG, "ttc_get_nhc: thermostat not set");
    std::vector<double> st(3 * c->nhc_M);
    TTC_HIP(hipStreamSynchronize(c->stream));
    TTC_HIP(hipMemcpy(st.data(), c->nhc, sizeof(double) * st.size(), hipMemcpyDeviceToHost));
    int M = c->nhc_M;
    for (int j = 0; j < M; j++) {
        zeta[3 * j] = st[j];
        zeta[3 * j + 1] = st[M + j];
        zeta[3 * j + 2] = st[2 * M + j];
    }
    return TTC_OK;
}

// nvt_md.c:528-531: integrate_nhc(dt); integrate_vverlet(k, 1.0); integrate_nhc(dt)
void ttc_launch_nhc_step(ttc_ctx *c, bool energy) {
    ttc_launch_kinet(c);
    ttc_launch_nhc(c);
    ttc_launch_vscale(c);
    ttc_launch_nhc_drift(c);
    ttc_side_fork(c, 0);
    {
        TtcOnSide side(c);
        ttc_launch_md_spread(c);
        ttc_launch_fft(c, energy);
    }
    ttc_launch_rebuild(c, true);
    ttc_side_fork(c, 1);
    {
        TtcOnSide side(c);
        ttc_launch_gather(c, energy, true, TTC_BT_ALL, false);
    }
    ttc_launch_pair_drifted(c, energy);
    ttc_side_join(c);
    ttc_launch_kick(c, 1.0, true);   // kinetic partials for the second thermostat half
    ttc_launch_nhc(c);
    ttc_launch_vscale(c);
    if (energy) {
        ttc_launch_kinet(c);
        ttc_launch_reduce(c, true);
    }
}

// The list format of the NHC loop: half lists (each pair once, partner
// forces through the windows) where the context is large enough, as in the
// velocity-Verlet loop.  The half-list format is made after the forces of
// the current positions, which the full list gives.
// These loops redo no step after an overflow (a run that overflows
// reports it at the end), so their half lists get margins first: the
// partner bitmap over the whole ring where that costs at most 16 KiB of
// LDS, and windows at least 30 % above the largest one the build at the
// anchors makes.
static int nhc_list(ttc_ctx *c, bool half) {
    if (c->half_mode != half) {
        c->half_mode = half;
        destroy_graphs(c);
        c->gen++;
    }
    int r = need_list(c, half);
    if (r) return r;
    if (!c->f4_valid) {
        launch_forces(c, false);
        c->f4_valid = true;
    }
    if (half && !c->list_half && (size_t)(c->n + 128) / 64 * 12 <= 16384 && c->span < c->n) {
        ttc_bitmap_sizes(c, c->n / 2 + 64, c->n / 2 + 64, &c->span, &c->span_pre);
        r = ttc_build_grp_attrs(c);
        if (r) return r;
        destroy_graphs(c);
        c->gen++;
    }
    r = ensure_format(c, half);
    if (r || !half) return r;
    const int wmax = c->h_status[ST_WMAX];
    if (wmax * 13 / 10 > c->wcap) {
        r = grow_window(c, wmax * 13 / 10, 0, 0);
        if (r) return r;
        c->list_half = false;   // build the half list again with the wider windows
        r = ensure_format(c, true);
    }
    return r;
}

// state a captured NHC step needs: list (half lists where `half`), row
// buffer, forces at the current positions
int ttc_prepare_nhc(ttc_ctx *c, int max_rows, bool half) {
    CHECK(c && c->topo && c->nhc_M > 0, TTC_ERR_ARG, "ttc_md_nhc: no topology or no thermostat");
    int r = ensure_rows(c, 1, max_rows);
    if (r) return r;
    r = nhc_list(c, half && half_enabled(c));
    if (r) return r;
    c->mts_valid = false;
    return TTC_OK;
}

extern "C" int ttc_md_nhc(ttc_ctx *c, int nsteps, int energy_every, double *energies, int max_rows, int *nrows,
                          int *rebuilds) {
    CHECK(c && c->topo && nsteps >= 0 && c->nhc_M > 0, TTC_ERR_ARG, "ttc_md_nhc: bad argument or no thermostat");
    CHECK(energy_every == 0 || (energies && max_rows > 0), TTC_ERR_ARG, "ttc_md_nhc: no row buffer");
    int r = ensure_rows(c, energy_every, max_rows);
    if (r) return r;
    r = nhc_list(c, half_enabled(c));
    if (r) return r;
    if (!c->g_nhc_plain) {
        r = capture_fn(c, &c->g_nhc_plain, [&] { ttc_launch_nhc_step(c, false); });
        if (r) return r;
        r = capture_fn(c, &c->g_nhc_energy, [&] { ttc_launch_nhc_step(c, true); });
        if (r) return r;
    }
    r = run_graphs(c, c->g_nhc_plain, c->g_nhc_energy, nsteps, energy_every, energies, max_rows, nrows, rebuilds);
    c->mts_valid = false;
    return r;
}

// ----------------------------------------------------------- MC energies

extern "C" int ttc_enbnd_mc(ttc_ctx *c, int lb, int ub, double *energy8) {
    CHECK(c && energy8 && lb >= 0 && ub <= c->n && lb <= ub, TTC_ERR_ARG, "ttc_enbnd_mc: bad argument");
    int r = need_list(c);
    if (r) return r;
    ttc_launch_pair_range(c, lb, ub);
    ttc_launch_spread(c);
    ttc_launch_pme(c, true, true, 0);
    ttc_launch_reduce(c, false);
    r = finish(c, nullptr, 0);
    if (r) return r;
    slots8(c, energy8);
    c->f4_valid = false;
    c->mts_valid = false;
    return TTC_OK;
}

extern "C" int ttc_nblist_pivot(ttc_ctx *c, int ub) {
    CHECK(c && ub >= 0 && ub <= c->n, TTC_ERR_ARG, "ttc_nblist_pivot: bad argument");
    return ttc_nblist(c);
}
