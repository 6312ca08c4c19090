// ttc_angles.hip -- angle-type bonded terms: bends (cbend.c:63-170),
// Urey-Bradley 1-3 pairs (curey.c:57-140), Fourier torsions
// (ctorsion.c:67-236) and harmonic impropers (cimproper.c:62-225), with the
// Hessian diagonal of each (cbend2.c, curey2.c, ctorsion2.c, cimproper2.c,
// CONFIGT).
//
// One thread per term, kinds in separate wave-aligned ranges so that a wave
// runs one code path.  A term's force on each of its sites goes to a slot of
// its own (no atomics); the gather that runs next adds a site's slots in a
// fixed CSR order, so the sum is the same every run.  Energies, virial and
// Hessian sums are reduced per block into the partials the reduce kernel
// sums in a fixed order.  FP64 throughout, the reference's formulas.
//
// Hessian diagonal: d2E/dx_i^2 for every coordinate x_i of a term's sites,
// by forward-mode second-order differentiation (value, first and second
// derivative along one coordinate, Taylor arithmetic) of the term's energy:
// nine (bend), six (Urey-Bradley) or twelve (dihedral) passes per term.
// Dihedrals are taken as atan2(sin, cos) there, which stays regular at 0
// and pi where the acos form's derivative does not.
#include "ttc_common.h"

namespace {

// value, first and second derivative along one coordinate
struct J2 {
    double v, d, dd;
};
__device__ __forceinline__ J2 jc(double v) { return {v, 0.0, 0.0}; }
__device__ __forceinline__ J2 operator+(J2 a, J2 b) { return {a.v + b.v, a.d + b.d, a.dd + b.dd}; }
__device__ __forceinline__ J2 operator-(J2 a, J2 b) { return {a.v - b.v, a.d - b.d, a.dd - b.dd}; }
__device__ __forceinline__ J2 operator-(J2 a) { return {-a.v, -a.d, -a.dd}; }
__device__ __forceinline__ J2 operator*(J2 a, J2 b) {
    return {a.v * b.v, a.v * b.d + a.d * b.v, a.v * b.dd + 2.0 * a.d * b.d + a.dd * b.v};
}
__device__ __forceinline__ J2 operator*(double s, J2 a) { return {s * a.v, s * a.d, s * a.dd}; }
// f(a) from f, f', f'' at a.v
__device__ __forceinline__ J2 jfun(J2 a, double f, double f1, double f2) {
    return {f, f1 * a.d, f1 * a.dd + f2 * a.d * a.d};
}
__device__ __forceinline__ J2 jsqrt(J2 a) {
    double s = sqrt(a.v);
    return jfun(a, s, 0.5 / s, -0.25 / (s * a.v));
}
__device__ __forceinline__ J2 jrcp(J2 a) {
    double r = 1.0 / a.v;
    return jfun(a, r, -r * r, 2.0 * r * r * r);
}
__device__ __forceinline__ J2 jacos(J2 a) {
    double q = 1.0 - a.v * a.v, s = sqrt(q);
    return jfun(a, acos(a.v), -1.0 / s, -a.v / (q * s));
}
__device__ __forceinline__ J2 jcos(J2 a) {
    double c = cos(a.v), s = sin(a.v);
    return jfun(a, c, -s, -c);
}
// atan2(y, x): d = (x dy - y dx) / r2, second derivative by the quotient rule
__device__ __forceinline__ J2 jatan2(J2 y, J2 x) {
    J2 r2 = x * x + y * y;
    double v = atan2(y.v, x.v);
    double n1 = x.v * y.d - y.v * x.d;                 // numerator of the first derivative
    double n1d = x.v * y.dd - y.v * x.dd;              // its derivative (the x'y' terms cancel)
    double d = n1 / r2.v;
    double dd = (n1d * r2.v - n1 * r2.d) / (r2.v * r2.v);
    return {v, d, dd};
}

template <typename T> struct V3 {
    T x, y, z;
};
template <typename T> __device__ __forceinline__ V3<T> vsub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> __device__ __forceinline__ T vdot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ V3<T> vcross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------- energies
// (the Hessian passes differentiate these; the forces below follow the
// reference's closed-form expressions)

// cbend.c: kb (phi - phi0)^2, phi the angle at b between b->a and b->c
__device__ __forceinline__ J2 bend_energy(V3<J2> u, V3<J2> w, double eq, double kb) {
    J2 c = vdot(u, w) * jrcp(jsqrt(vdot(u, u)) * jsqrt(vdot(w, w)));
    J2 dphi = jacos(c) - jc(eq);
    return kb * (dphi * dphi);
}

// curey.c: k (r - S0)^2
__device__ __forceinline__ J2 urey_energy(V3<J2> d, double s0, double k) {
    J2 dr = jsqrt(vdot(d, d)) - jc(s0);
    return k * (dr * dr);
}

// dihedral of the bond vectors b1 = b - a, b2 = c - b, b3 = d - c, with the
// sign convention of ctorsion.c:171-182
__device__ __forceinline__ J2 dihedral_j(V3<J2> b1, V3<J2> b2, V3<J2> b3) {
    V3<J2> n1 = vcross(b1, b2), n2 = vcross(b2, b3);
    V3<J2> m = vcross(n1, n2);
    J2 y = vdot(m, b2) * jrcp(jsqrt(vdot(b2, b2)));
    return jatan2(y, vdot(n1, n2));
}

// ---------------------------------------------------------------- kernel

struct AngPack {
    int ntot;                  // padded terms
    int st1, st2, st3;         // first term of urey, tors, impr (bend at 0)
    int sl1, sl2, sl3;         // first force slot of urey, tors, impr
    const int4 *idx;
    const double2 *par;
    const double *tpar;
    const double4 *pos;        // sorted order through inv, or original order (inv null)
    const int *inv;
    double lx, ly, lz, hx, hy, hz;
    double4 *tf;
    double *partials;
    int maxb;
    int kinds;                 // bit k: kind k evaluated
};

__device__ __forceinline__ V3<double> site(const AngPack &a, int i) {
    double4 p = a.pos[a.inv ? a.inv[i] : i];
    return {p.x, p.y, p.z};
}
__device__ __forceinline__ V3<double> mi3(const AngPack &a, V3<double> d) {
    return {ttc_minimg(d.x, a.hx, a.lx), ttc_minimg(d.y, a.hy, a.ly), ttc_minimg(d.z, a.hz, a.lz)};
}
__device__ __forceinline__ void put(double4 *tf, int slot, double fx, double fy, double fz) {
    tf[slot] = make_double4(fx, fy, fz, 0.0);
}
__device__ __forceinline__ V3<J2> jv(V3<double> d, int axis, double seed) {
    V3<J2> r = {jc(d.x), jc(d.y), jc(d.z)};
    if (axis == 0) r.x.d = seed;
    else if (axis == 1) r.y.d = seed;
    else r.z.d = seed;
    return r;
}

// virial of a dihedral term (ctorsion.c:252-261)
__device__ __forceinline__ void dihedral_virial(const double fA[3], const double fC[3], const double fD[3],
                                                const double d1[3], const double d2[3], const double d3[3],
                                                double *v) {
    const int ii[6] = {0, 1, 1, 2, 2, 2}, jj[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        int i = ii[q], j = jj[q];
        v[q] = -fA[i] * d1[j] + fC[i] * d2[j] + fD[i] * (d2[j] + d3[j]);
    }
}

template <bool ENERGY, bool HES>
__device__ __forceinline__ void angles_body(const AngPack &a, int bid) {
    int t = bid * blockDim.x + threadIdx.x;
    double ek[4] = {0.0, 0.0, 0.0, 0.0}, vir[6] = {0, 0, 0, 0, 0, 0}, hs[3] = {0.0, 0.0, 0.0};
    if (t < a.ntot) {
        int kind = (t >= a.st1) + (t >= a.st2) + (t >= a.st3);
        int4 q = a.idx[t];
        if (q.x >= 0 && ((a.kinds >> kind) & 1)) {
            double2 pr = a.par[t];
            if (kind == 0) {
                // ---- bend (cbend.c:63-170)
                int loc = t;
                V3<double> pa = site(a, q.x), pb = site(a, q.y), pc = site(a, q.z);
                V3<double> u = mi3(a, vsub(pa, pb)), w = mi3(a, vsub(pc, pb));
                double iu = 1.0 / sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
                double eux = u.x * iu, euy = u.y * iu, euz = u.z * iu;
                double iw = 1.0 / sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
                double ewx = w.x * iw, ewy = w.y * iw, ewz = w.z * iw;
                double cs = eux * ewx + euy * ewy + euz * ewz;
                double phi = acos(cs);
                double isn = 1.0 / sqrt(1.0 - cs * cs);
                double dphi = phi - pr.x;
                double fr = -pr.y * dphi * 2.0;
                double ga = iu * isn * fr, gc = iw * isn * fr;
                double fax = (eux * cs - ewx) * ga, fay = (euy * cs - ewy) * ga, faz = (euz * cs - ewz) * ga;
                double fcx = (ewx * cs - eux) * gc, fcy = (ewy * cs - euy) * gc, fcz = (ewz * cs - euz) * gc;
                put(a.tf, 3 * loc, fax, fay, faz);
                put(a.tf, 3 * loc + 1, -(fax + fcx), -(fay + fcy), -(faz + fcz));
                put(a.tf, 3 * loc + 2, fcx, fcy, fcz);
                if (ENERGY) {
                    ek[0] = -fr * dphi * 0.5;
                    vir[0] = fax * u.x + fcx * w.x; vir[1] = fay * u.x + fcy * w.x; vir[2] = fay * u.y + fcy * w.y;
                    vir[3] = faz * u.x + fcz * w.x; vir[4] = faz * u.y + fcz * w.y; vir[5] = faz * u.z + fcz * w.z;
                }
                if (HES) {
                    // site a moves u, site c moves w, site b moves both by -1
                    for (int ax = 0; ax < 3; ax++) {
                        V3<J2> U0 = jv(u, ax, 0.0), W0 = jv(w, ax, 0.0);
                        hs[ax] += bend_energy(jv(u, ax, 1.0), W0, pr.x, pr.y).dd;
                        hs[ax] += bend_energy(U0, jv(w, ax, 1.0), pr.x, pr.y).dd;
                        hs[ax] += bend_energy(jv(u, ax, -1.0), jv(w, ax, -1.0), pr.x, pr.y).dd;
                    }
                }
            } else if (kind == 1) {
                // ---- Urey-Bradley pair (curey.c:57-140)
                int loc = t - a.st1;
                V3<double> d = mi3(a, vsub(site(a, q.y), site(a, q.x)));
                double dr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z), dri = 1.0 / dr;
                double rr = dr - pr.x;
                double fr = -pr.y * rr * 2.0;
                double fx = fr * dri * d.x, fy = fr * dri * d.y, fz = fr * dri * d.z;
                put(a.tf, a.sl1 + 2 * loc, -fx, -fy, -fz);
                put(a.tf, a.sl1 + 2 * loc + 1, fx, fy, fz);
                if (ENERGY) {
                    ek[1] = -(fr * rr * 0.5);
                    vir[0] = fx * d.x; vir[1] = fy * d.x; vir[2] = fy * d.y;
                    vir[3] = fz * d.x; vir[4] = fz * d.y; vir[5] = fz * d.z;
                }
                if (HES)
                    for (int ax = 0; ax < 3; ax++)
                        hs[ax] += 2.0 * urey_energy(jv(d, ax, 1.0), pr.x, pr.y).dd;
            } else {
                // ---- dihedral: Fourier torsion (ctorsion.c:67-236) or
                // harmonic improper (cimproper.c:62-225)
                bool tors = kind == 2;
                int loc = t - (tors ? a.st2 : a.st3);
                int sl = (tors ? a.sl2 : a.sl3) + 4 * loc;
                V3<double> P0 = site(a, q.x), P1 = site(a, q.y), P2 = site(a, q.z), P3 = site(a, q.w);
                V3<double> b1 = mi3(a, vsub(P1, P0)), b2 = mi3(a, vsub(P2, P1)), b3 = mi3(a, vsub(P3, P2));
                double r1 = sqrt(b1.x * b1.x + b1.y * b1.y + b1.z * b1.z), i1 = 1.0 / r1;
                double r2 = sqrt(b2.x * b2.x + b2.y * b2.y + b2.z * b2.z), i2 = 1.0 / r2;
                double r3 = sqrt(b3.x * b3.x + b3.y * b3.y + b3.z * b3.z), i3 = 1.0 / r3;
                double e1[3] = {b1.x * i1, b1.y * i1, b1.z * i1};
                double e2[3] = {b2.x * i2, b2.y * i2, b2.z * i2};
                double e3[3] = {b3.x * i3, b3.y * i3, b3.z * i3};
                double n1[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
                double cb = -(e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]);
                double isb2 = 1.0 / (1.0 - cb * cb), isb = sqrt(isb2);
                double n2[3] = {e3[1] * e2[2] - e3[2] * e2[1], e3[2] * e2[0] - e3[0] * e2[2], e3[0] * e2[1] - e3[1] * e2[0]};
                double cc = -(e3[0] * e2[0] + e3[1] * e2[1] + e3[2] * e2[2]);
                double isc2 = 1.0 / (1.0 - cc * cc), isc = sqrt(isc2);
                double mx = -(n1[1] * n2[2] - n1[2] * n2[1]);
                double my = -(n1[2] * n2[0] - n1[0] * n2[2]);
                double mz = -(n1[0] * n2[1] - n1[1] * n2[0]);
                double sgn = (mx * e2[0] + my * e2[1] + mz * e2[2]) >= 0.0 ? 1.0 : -1.0;
                double ct = -(n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) * isb * isc;
                ct = fmin(1.0, fmax(-1.0, ct));
                double tau = sgn * acos(ct);
                double fr = 0.0, e = 0.0;
                const double *tp = a.tpar + (size_t)loc * 18;
                if (tors) {
                    for (int m = 0; m < 6; m++) {
                        double kk = tp[m];
                        if (kk == 0.0) break;
                        double arg = tp[6 + m] * tau - tp[12 + m];
                        fr += kk * tp[6 + m] * sin(arg);
                        e += kk * (1 + cos(arg));
                    }
                } else {
                    double dt = tau - pr.x;
                    fr = -2.0 * pr.y * dt;
                    e = -fr * dt * 0.5;
                }
                double fa = -fr * i1 * isb2;
                double fb1 = fr * (r2 - r1 * cb) * i1 * i2 * isb2, fb2 = fr * cc * i2 * isc2;
                double fc1 = fr * (r2 - r3 * cc) * i3 * i2 * isc2, fc2 = fr * cb * i2 * isb2;
                double fd = -fr * i3 * isc2;
                double fA[3], fB[3], fC[3], fD[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    fA[k] = fa * n1[k];
                    fB[k] = fb1 * n1[k] + fb2 * n2[k];
                    fC[k] = fc1 * n2[k] + fc2 * n1[k];
                    fD[k] = fd * n2[k];
                }
                put(a.tf, sl, fA[0], fA[1], fA[2]);
                put(a.tf, sl + 1, fB[0], fB[1], fB[2]);
                put(a.tf, sl + 2, fC[0], fC[1], fC[2]);
                put(a.tf, sl + 3, fD[0], fD[1], fD[2]);
                if (ENERGY) {
                    ek[kind] = e;
                    double d1[3] = {b1.x, b1.y, b1.z}, d2[3] = {b2.x, b2.y, b2.z}, d3[3] = {b3.x, b3.y, b3.z};
                    dihedral_virial(fA, fC, fD, d1, d2, d3, vir);
                }
                if (HES) {
                    // site k moves b_k by +1 and b_{k+1} by -1 (b_0, b_4 absent)
                    for (int s = 0; s < 4; s++)
                        for (int ax = 0; ax < 3; ax++) {
                            V3<J2> B1 = jv(b1, ax, s == 1 ? 1.0 : (s == 0 ? -1.0 : 0.0));
                            V3<J2> B2 = jv(b2, ax, s == 2 ? 1.0 : (s == 1 ? -1.0 : 0.0));
                            V3<J2> B3 = jv(b3, ax, s == 3 ? 1.0 : (s == 2 ? -1.0 : 0.0));
                            J2 th = dihedral_j(B1, B2, B3);
                            J2 E;
                            if (tors) {
                                E = jc(0.0);
                                for (int m = 0; m < 6; m++) {
                                    double kk = tp[m];
                                    if (kk == 0.0) break;
                                    E = E + kk * jcos(tp[6 + m] * th - jc(tp[12 + m]));
                                }
                            } else {
                                // the improper's (tau - tau0) with tau on the
                                // branch of the value the force used
                                J2 dt = th - jc(pr.x);
                                dt.v = tau - pr.x;
                                E = pr.y * (dt * dt);
                            }
                            hs[ax] += E.dd;
                        }
                }
            }
        }
    }
    if (!ENERGY) return;
    // per-block sums: energies of the evaluated kinds, virial, Hessian
    __shared__ double red[4][13];
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double vals[13] = {ek[0], ek[1], ek[2], ek[3], vir[0], vir[1], vir[2], vir[3], vir[4], vir[5], hs[0], hs[1], hs[2]};
    const int nv = HES ? 13 : 10;
    for (int k = 0; k < nv; k++) {
        double s = ttc_wave_sum(vals[k]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        int k = threadIdx.x;
        double s = red[0][k] + red[1][k] + red[2][k] + red[3][k];
        int slot = k < 4 ? EV_EBEND + k : (k < 10 ? EV_AV0 + (k - 4) : EV_AHX + (k - 10));
        // a kind this launch did not evaluate keeps the energy another launch left
        if (k >= 4 || ((a.kinds >> k) & 1)) a.partials[(size_t)slot * a.maxb + bid] = s;
    }
}

template <bool ENERGY, bool HES>
__global__ void __launch_bounds__(256) k_angles(AngPack a) { angles_body<ENERGY, HES>(a, blockIdx.x); }

template <bool ENERGY>
__global__ void __launch_bounds__(256) kb_angles(const AngPack *pk, const TtcBmap *map) {
    const TtcBmap m = map[blockIdx.x];
    const AngPack a = pk[m.rep];
    angles_body<ENERGY, false>(a, m.blk);
}

AngPack ang_pack(const ttc_ctx *c, int bterms, bool snap) {
    AngPack a;
    a.ntot = c->ang_start[4];
    a.st1 = c->ang_start[1]; a.st2 = c->ang_start[2]; a.st3 = c->ang_start[3];
    a.sl1 = c->ang_slot[1]; a.sl2 = c->ang_slot[2]; a.sl3 = c->ang_slot[3];
    a.idx = c->ang_idx; a.par = c->ang_par; a.tpar = c->ang_tpar;
    a.pos = snap ? c->snap : c->pos4;
    a.inv = snap ? nullptr : c->inv;
    a.lx = c->p.lx; a.ly = c->p.ly; a.lz = c->p.lz;
    a.hx = 0.5 * c->p.lx; a.hy = 0.5 * c->p.ly; a.hz = 0.5 * c->p.lz;
    a.tf = c->ang_tf;
    a.partials = c->partials;
    a.maxb = c->nblk_max;
    a.kinds = (bterms >> 1) & 15;
    return a;
}

}  // namespace

int ttc_angle_blocks(const ttc_ctx *c) { return (c->ang_start[4] + 255) / 256; }

void ttc_launch_angles(ttc_ctx *c, bool energy, bool hes, int bterms, bool snap) {
    if (!c->ang_start[4] || !((bterms >> 1) & 15)) return;
    AngPack a = ang_pack(c, bterms, snap);
    int blocks = ttc_angle_blocks(c);
    if (hes) k_angles<true, true><<<blocks, 256, 0, c->stream>>>(a);
    else if (energy) k_angles<true, false><<<blocks, 256, 0, c->stream>>>(a);
    else k_angles<false, false><<<blocks, 256, 0, c->stream>>>(a);
}

// batched launches (replica batches): every replica's terms in one launch
int ttc_bstage_angles(const std::vector<ttc_ctx *> &r, TtcStage *st) {
    std::vector<AngPack> pk;
    std::vector<int> nb;
    for (ttc_ctx *c : r) {
        pk.push_back(ang_pack(c, TTC_BT_ALL, false));
        nb.push_back(ttc_angle_blocks(c));
    }
    return ttc_stage_make(st, pk, nb, 0);
}

void ttc_blaunch_angles(const TtcStage &st, bool energy, hipStream_t s) {
    if (!st.nblk) return;
    if (energy) kb_angles<true><<<st.nblk, 256, 0, s>>>((const AngPack *)st.pk, st.map);
    else kb_angles<false><<<st.nblk, 256, 0, s>>>((const AngPack *)st.pk, st.map);
}
