"""CPU pins of the oracle's sampler (no GPU).

  * splitmix64 and xoshiro256** against their published output: the first
    ten xoshiro256** outputs from the state {1, 2, 3, 4} and the first four
    splitmix64 outputs from 0, as the authors' reference code prints them;
  * orc_dos_interp bit-exact against the reference's own dos_interp
    (dos_interpolate.c:107-124, DOS build, compiled into oracle/_ref);
  * orc_wl_run (pivot.c:236-253 acceptance, mcmove_pivot.c / mcmove_displ.c
    moves) on a 4-bead chain with bond energy only, whose density of states
    is closed (tests/sampling_exact.py):
      - at fixed ln g = beta E_mid the walk samples exp(-beta E) exactly
        (interpolated ln g of a linear function is that function), so the
        bin histogram matches the Boltzmann integral of Omega(E); the same
        data reject the prediction of an acceptance without interpolation;
      - with the end-to-end repulsion switched on (sig_rep 3 A) the same
        histogram matches an independent importance-sampled estimate;
      - a Wang-Landau run converges to the ln g at which interpolated
        acceptance visits every bin equally.
"""
import numpy as np
import pytest

import pyoracle
import sampling_exact as X

L = pyoracle.lib()
needs_dos_ref = pytest.mark.skipif(not __import__("os").path.exists(pyoracle.REF_DOS_SO),
                                   reason="oracle/_ref not built (no reference tree)")

# Blackman and Vigna's reference xoshiro256starstar.c from s = {1, 2, 3, 4}
XOSHIRO_1234 = [11520, 0, 1509978240, 1215971899390074240, 1216172134540287360, 607988272756665600,
                16172922978634559625, 8476171486693032832, 10595114339597558777, 2904607092377533576]
# Vigna's splitmix64.c from x = 0
SPLITMIX_0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_xoshiro_published_output():
    s = np.array([1, 2, 3, 4], dtype=np.uint64)
    got = [int(L.orc_xoshiro_next(pyoracle._u64(s))) for _ in range(10)]
    assert got == XOSHIRO_1234


def test_splitmix_published_output():
    assert [int(v) for v in pyoracle.splitmix_seq(0, 4)] == SPLITMIX_0


def test_xoshiro_uniform_is_top_53_bits():
    s = np.array([5, 6, 7, 8], dtype=np.uint64)
    s2 = s.copy()
    for _ in range(100):
        u = L.orc_xoshiro_uniform(pyoracle._u64(s))
        r = int(L.orc_xoshiro_next(pyoracle._u64(s2)))
        assert u == (r >> 11) * 2.0 ** -53 and 0.0 <= u < 1.0


@needs_dos_ref
def test_dos_interp_bit_exact_vs_reference():
    R = pyoracle.ref_dos_lib()
    rng = np.random.default_rng(11)
    for nbins in (2, 3, 7, 64, 1000):
        for _ in range(40):
            lng = np.cumsum(rng.normal(0.0, 3.0, nbins)) + rng.uniform(-50, 50)
            e0, w = rng.uniform(-500, 500), rng.uniform(0.05, 8.0)
            for b in {0, 1 % nbins, nbins - 1, int(rng.integers(0, nbins))}:
                for val in (e0 + w * b, e0 + w * (b + 0.5), e0 + w * (b + rng.uniform()), e0 + w * (b + 1) * (1 - 1e-16)):
                    ref = R.ref_dos_interp(nbins, lng, e0, w, b, val)
                    got = L.orc_dos_interp(nbins, lng, e0, w, b, val)
                    assert got == ref, (nbins, b, val, got, ref)


def _params(**kw):
    d = dict(kbond=X.KBOND, r0bond=X.R0, eps=1.0, sig_rep=0.01, pivot_frac=0.5, theta_max=0.5, disp_max=1.5,
             emin=0.0, ebin=X.EBIN, nbins=X.NBINS)
    d.update(kw)
    return pyoracle.WlParams(**d)


class _Walker:
    """One 4-bead oracle walker whose state persists across runs."""

    def __init__(self, p, seed):
        self.p = p
        self.nat = np.zeros((4, 4), np.uint8)
        self.r0m = np.zeros((4, 4))
        self.pos = X.start_positions()
        self.e = 3 * X.KBOND * 0.3 ** 2
        self.rng = pyoracle.splitmix_seq(seed, 256)
        self.stats = np.zeros(7, np.int64)

    def run(self, lng, ntrials, lnf):
        hist = np.zeros(self.p.nbins, np.int64)
        self.e = pyoracle.wl_run(self.p, self.pos, self.e, self.nat, self.r0m, lng, hist, self.rng, ntrials, lnf,
                                 self.stats)
        return hist

    def energy(self):
        return L.orc_go_energy(4, self.pos.ravel().copy(), X.KBOND, X.R0, 0, np.zeros(1, np.int32),
                               np.zeros(1, np.int32), np.zeros(1), self.p.eps, self.p.sig_rep, 1e30)


def test_wl_start_energy_is_the_chain_energy():
    w = _Walker(_params(), 1)
    assert abs(w.energy() - w.e) < 1e-12


def test_wl_fixed_g_samples_boltzmann_exactly():
    beta = 1.0
    p = _params()
    mids = (np.arange(X.NBINS) + 0.5) * X.EBIN
    fr = []
    for sd in range(24):
        w = _Walker(p, 1 + sd)
        lng = beta * mids
        hist = w.run(lng, 250000, 0.0)
        st = w.stats
        assert st[0] == 250000 and st[3] > 0.8 * st[0] and 0.3 < st[1] / st[0] < 0.95
        assert np.array_equal(lng, beta * mids) and hist.sum() == 250000
        assert abs(w.energy() - w.e) < 1e-9     # the running energy is the chain's
        fr.append(hist / hist.sum())
    fr = np.array(fr)
    m = fr.mean(0)
    se = fr.std(0, ddof=1) / np.sqrt(len(fr))
    exact = X.boltzmann_visits(beta)
    z = (m - exact) / se
    print("z-scores vs exact:", np.round(z, 2))
    assert np.abs(z).max() < 4.5
    assert (z ** 2).sum() < 45.0      # chi^2, 16 bins (p ~ 1.5e-4)
    # acceptance with ln g of the bins instead of interpolated: rejected
    stepwise = np.diff(X.omega_cum(np.arange(X.NBINS + 1) * X.EBIN)) * np.exp(-beta * mids)
    stepwise /= stepwise.sum()
    assert np.abs((m - stepwise) / se).max() > 8.0


def test_wl_fixed_g_with_repulsion_matches_independent_sampling():
    beta = 1.0
    p = _params(sig_rep=3.0)
    mids = (np.arange(X.NBINS) + 0.5) * X.EBIN
    fr = []
    for sd in range(24):
        w = _Walker(p, 300 + sd)
        w.e = w.energy()
        w.run(beta * mids, 250000, 0.0)
        hist = w.run(beta * mids, 250000, 0.0)    # the first run equilibrates
        assert abs(w.energy() - w.e) < 1e-9
        fr.append(hist / hist.sum())
    fr = np.array(fr)
    m = fr.mean(0)
    se = fr.std(0, ddof=1) / np.sqrt(len(fr))
    ref, rse = X.boltzmann_visits_sampled(beta, 3.0, 1.0, 2_000_000, seed=5)
    z = (m - ref) / np.sqrt(se ** 2 + rse ** 2)
    print("z-scores vs independent sampling:", np.round(z, 2))
    assert np.abs(z).max() < 4.5
    # the repulsion moves the histogram far beyond these errors
    assert np.abs((m - X.boltzmann_visits(beta)) / se).max() > 10.0


def test_wl_converges_to_interpolated_fixed_point():
    p = _params()
    fp = X.wl_fixed_point()
    res = []
    for sd in range(8):
        w = _Walker(p, 1000 + sd)
        lng = np.zeros(X.NBINS)
        lnf = 0.1
        while lnf > 1e-4:
            w.run(lng, 200000, lnf)
            lnf /= 2
        assert abs(w.energy() - w.e) < 1e-9
        res.append(lng - lng.mean())
    m = np.array(res).mean(0)
    d = m - fp
    print("ln g - fixed point:", np.round(d, 4), "rms", np.sqrt((d ** 2).mean()))
    # a Wang-Landau run saturates at an error set by its last ln f and length
    # (about 0.05 per run here); eight runs averaged
    assert np.abs(d).max() < 0.04 and np.sqrt((d ** 2).mean()) < 0.02
    # the binned ln g without the interpolation's correction is farther off
    # in the first bin than the run's error there
    lgb = X.ln_g_bins()
    lgb = lgb - lgb.mean()
    assert abs(fp[0] - lgb[0]) > 0.04
