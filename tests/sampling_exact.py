"""Exact distributions of a 4-bead Go walker with bond energy only.

Three bonds of energy kbond (r - r0)^2 each, no pair term (the one pair with
|i - j| >= 3 carries a repulsion eps (sig_rep / r)^12 with sig_rep small
enough to vanish at every reachable distance).  Positions are sampled with
uniform measure d^3x per bead, i.e. d^3b per bond vector, r^2 dr per bond
length.  With u_i = r_i - r0 and rho^2 = sum u_i^2 = E / kbond, the density
of states is

    Omega(E) ~ rho * < prod_i (r0 + u_i)^2 >_{|u| = rho}
             = rho * (r0^6 + r0^4 rho^2 + r0^2 rho^4 / 5 + rho^6 / 105)

(moments of the uniform sphere in 3 dimensions: <u_i^2> = rho^2 / 3,
<u_i^2 u_j^2> = rho^4 / 15, <u_1^2 u_2^2 u_3^2> = rho^6 / 105; odd moments
vanish), valid while rho < r0.  Its integral over [E_a, E_b] is closed:
2 kbond [r0^6 rho^3 / 3 + r0^4 rho^5 / 5 + r0^2 rho^7 / 35 + rho^9 / 945].
"""
import numpy as np
from scipy.integrate import quad

KBOND = 1.0
R0 = 3.8
NBINS = 16
EBIN = 0.5


def omega(e, k=KBOND, r0=R0):
    rho = np.sqrt(e / k)
    return rho * (r0 ** 6 + r0 ** 4 * rho ** 2 + r0 ** 2 * rho ** 4 / 5 + rho ** 6 / 105)


def omega_cum(e, k=KBOND, r0=R0):
    rho = np.sqrt(e / k)
    return 2 * k * (r0 ** 6 * rho ** 3 / 3 + r0 ** 4 * rho ** 5 / 5 + r0 ** 2 * rho ** 7 / 35 + rho ** 9 / 945)


def ln_g_bins(nbins=NBINS, ebin=EBIN):
    """ln of the states in each bin [b w, (b + 1) w), closed form."""
    edges = np.arange(nbins + 1) * ebin
    return np.log(np.diff(omega_cum(edges)))


def interp(lng, e, ebin=EBIN, emin=0.0):
    """ln g at e as dos_interpolate.c:107-124 interpolates it."""
    nb = len(lng)
    b = min(int((e - emin) / ebin), nb - 1)
    mid = emin + ebin * (b + 0.5)
    if b == 0:
        s = (lng[1] - lng[0]) / ebin
    elif b == nb - 1:
        s = (lng[b] - lng[b - 1]) / ebin
    else:
        s = (lng[b + 1] - lng[b - 1]) * 0.5 / ebin
    return lng[b] + s * (e - mid)


def visits(lng, ebin=EBIN):
    """Stationary fraction of trials ending in each bin when the walk samples
    exp(-interpolated ln g(E)) (fixed g): int_bin Omega(E) exp(-G(E)) dE."""
    nb = len(lng)
    ref = lng.mean()
    h = np.array([quad(lambda e: omega(e) * np.exp(-(interp(lng, e, ebin) - ref)), b * ebin, (b + 1) * ebin,
                       epsabs=0, epsrel=1e-11)[0] for b in range(nb)])
    return h / h.sum()


def boltzmann_visits(beta, nbins=NBINS, ebin=EBIN):
    """Bin fractions of exp(-beta E) sampling inside the window."""
    h = np.array([quad(lambda e: omega(e) * np.exp(-beta * e), b * ebin, (b + 1) * ebin,
                       epsabs=0, epsrel=1e-11)[0] for b in range(nbins)])
    return h / h.sum()


def wl_fixed_point(nbins=NBINS, ebin=EBIN, iters=200):
    """The ln g (up to a constant) at which interpolated Wang-Landau
    acceptance visits every bin equally: the limit of a converged run."""
    lng = ln_g_bins(nbins, ebin)
    for _ in range(iters):
        h = visits(lng, ebin)
        step = np.log(h * nbins)
        lng = lng + step
        if np.abs(step).max() < 1e-12:
            break
    return lng - lng.mean()


def start_positions(r0=R0):
    """A straight 4-bead chain along x with bonds 0.3 A longer than r0 (E > 0)."""
    pos = np.zeros((4, 3))
    pos[:, 0] = np.arange(4) * (r0 + 0.3)
    return pos


def boltzmann_visits_sampled(beta, sig_rep, eps, nsamp=4_000_000, seed=0, nbins=NBINS, ebin=EBIN,
                             k=KBOND, r0=R0):
    """Bin fractions (and their standard errors) of exp(-beta E) inside the
    window for the 4-bead chain with the end-to-end repulsion
    eps (sig_rep / r_03)^12 switched on, by independent sampling: bond
    lengths drawn from r^2 exp(-beta k (r - r0)^2) (rejection from the
    normal), directions uniform, each sample weighted by exp(-beta E_rep).
    An unbiased estimate that shares nothing with the Markov chain."""
    rng = np.random.default_rng(seed)
    sd = 1.0 / np.sqrt(2.0 * beta * k)
    rmax = r0 + 8.0 * sd
    acc = np.zeros(nbins)
    acc2 = np.zeros(nbins)
    chunk = 500_000
    done = 0
    while done < nsamp:
        m = min(chunk, nsamp - done)
        bonds = []
        for _ in range(3):
            r = np.empty(0)
            while r.size < m:
                c = rng.normal(r0, sd, 2 * m)
                c = c[(c > 0) & (rng.uniform(size=c.size) < (c / rmax) ** 2)]
                r = np.concatenate([r, c])
            r = r[:m]
            v = rng.normal(size=(m, 3))
            v /= np.linalg.norm(v, axis=1)[:, None]
            bonds.append((r, v * r[:, None]))
        eb = sum(k * (r - r0) ** 2 for r, _ in bonds)
        r03 = np.linalg.norm(bonds[0][1] + bonds[1][1] + bonds[2][1], axis=1)
        erep = eps * (sig_rep / r03) ** 12
        e = eb + erep
        wgt = np.exp(-beta * erep)
        ok = np.isfinite(e) & (e < nbins * ebin)
        b = np.where(ok, np.floor(np.where(ok, e, 0.0) / ebin), -1).astype(np.int64)
        ok &= (b >= 0) & (b < nbins)
        acc += np.bincount(b[ok], weights=wgt[ok], minlength=nbins)
        acc2 += np.bincount(b[ok], weights=wgt[ok] ** 2, minlength=nbins)
        done += m
    tot = acc.sum()
    frac = acc / tot
    # delta-method standard error of a ratio of weighted sums
    se = np.sqrt(np.maximum(acc2 * (1 - frac) ** 2 + (acc2.sum() - acc2) * frac ** 2, 0.0)) / tot
    return frac, se
