"""Pins of the OU input oracle (oracle/inputs.py)."""
import numpy as np
import scipy.stats

from oracle import inputs as OI

M64 = (1 << 64) - 1


def _mix64_int(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_counter_normal_integer_stream():
    """The uniforms come from the same 64-bit words a Python-integer
    restatement gives; splitmix64 of 0 + G is the published first output
    0xE220A8397B1DCDAF."""
    G = 0x9E3779B97F4A7C15
    assert _mix64_int(G) == 0xE220A8397B1DCDAF
    seed, j, n = 123456789, 7, 4242
    key = _mix64_int((seed * G + j) & M64)
    h1 = _mix64_int((key + 2 * n * G) & M64)
    h2 = _mix64_int((key + (2 * n + 1) * G) & M64)
    u1 = ((h1 >> 11) + 0.5) * 2.0 ** -53
    u2 = (h2 >> 11) * 2.0 ** -53
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
    assert OI.counter_normal(seed, [j], [n])[0, 0] == z


def test_counter_normal_distribution():
    """Standard normal: mean, variance, KS distance over 2^20 draws; streams
    and steps are uncorrelated."""
    z = OI.counter_normal(3, np.arange(64), np.arange(1 << 14))
    x = z.ravel()
    assert abs(x.mean()) < 5 / np.sqrt(x.size)
    assert abs(x.var() - 1.0) < 5 * np.sqrt(2.0 / x.size)
    assert scipy.stats.kstest(x, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(z[:, :-1].ravel(), z[:, 1:].ravel())[0, 1]) < 5 / np.sqrt(x.size)
    assert abs(np.corrcoef(z[:-1].ravel(), z[1:].ravel())[0, 1]) < 5 / np.sqrt(x.size)


def test_ou_process_statistics_and_update():
    """Stationary OU: mean mu, variance sigma^2, lag-k autocorrelation
    exp(-k dt/tau); consecutive values give back the normals exactly up to
    rounding (a wrong decay or noise scale fails this)."""
    dt, mu, sigma, tau = 1e-4, 0.5e-9, 0.2e-9, 5e-3
    I = OI.ou_process(32, 20000, dt, mu, sigma, tau, seed=11)
    a = np.exp(-dt / tau)
    y = (I - mu) / sigma
    neff = y.size * (1 - a) / (1 + a)           # effective samples of a correlated series
    assert abs(y.mean()) < 5 / np.sqrt(neff)
    assert abs(y.var() - 1.0) < 0.05
    for k in (1, 10, 50):
        r = np.mean(y[:, :-k] * y[:, k:]) / y.var()
        assert abs(r - a ** k) < 0.05, (k, r, a ** k)
    z = OI.counter_normal(11, np.arange(32), np.arange(20000))
    b = np.sqrt(1 - a * a)
    zr = (y[:, 1:] - a * y[:, :-1]) / b
    np.testing.assert_allclose(zr, z[:, 1:], atol=1e-9)
    np.testing.assert_allclose(y[:, 0], z[:, 0], rtol=0, atol=1e-12)


def test_binomial_sampler_pins():
    """The inverse-CDF binomial sampler (PoissonInput): its CDF table
    equals scipy's Binomial CDF, a brute-force enumeration of 2^3 outcomes
    for n = 3, the draws follow the pmf (chi-square), and p = 0 / p = 1
    give 0 / n."""
    from itertools import product
    from oracle import brian1 as O
    for n, p in ((10, 0.005), (1000, 0.002), (50, 0.3), (3, 0.25)):
        np.testing.assert_allclose(O.binomial_cdf_table(n, p), scipy.stats.binom.cdf(np.arange(n + 1), n, p),
                                   rtol=1e-12, atol=1e-15)
    pmf = np.zeros(4)
    for outcome in product((0, 1), repeat=3):
        pmf[sum(outcome)] += np.prod([0.25 if o else 0.75 for o in outcome])
    np.testing.assert_allclose(O.binomial_cdf_table(3, 0.25), np.cumsum(pmf), rtol=1e-15)
    bits = O.counter_u64(17, np.arange(64), np.arange(4096)[:, None]).ravel()
    k = O.binomial_inv(bits, 20, 0.1)
    obs = np.bincount(k, minlength=21)[:21]
    exp = scipy.stats.binom.pmf(np.arange(21), 20, 0.1) * len(k)
    m = exp > 5
    chi2 = ((obs[m] - exp[m]) ** 2 / exp[m]).sum() + (obs[~m].sum() - exp[~m].sum()) ** 2 / exp[~m].sum()
    assert scipy.stats.chi2.sf(chi2, m.sum()) > 1e-3
    assert (O.binomial_inv(bits[:10], 5, 0.0) == 0).all() and (O.binomial_inv(bits[:10], 5, 1.0) == 5).all()


def test_coincidence_counter_restatements_agree():
    """The event-driven CoincidenceCounter (monitor.py:1336-1376) and the
    per-step GPU runsim counter (gpu_modelfitting.py:118-156) are two
    formulations of one exclusive algorithm: equal counts on random trains;
    hand cases: a model spike within delta of one target spike counts once,
    one between two close targets counts once, onset excludes early spikes,
    target_length counts the target spikes passed."""
    from oracle import fitting as F
    rng = np.random.default_rng(3)
    BIG = F.BIG
    for _ in range(40):
        tgt = np.unique(rng.integers(50, 5000, rng.integers(5, 60)))
        mod = np.unique(rng.integers(50, 5000, rng.integers(5, 60)))
        data = np.concatenate(([-BIG], tgt, [BIG]))
        c, m, tl = F.coincidence_counter([mod], data, [0], 40, 1e-4)
        assert c[0] == F.exclusive_coincidences_single(mod, tgt, 40)
        assert m[0] == len(mod)
        assert tl[0] == np.sum(tgt < mod.max())
    data = np.array([-BIG, 100, 130, 1000, BIG])
    c, m, tl = F.coincidence_counter([np.array([120])], data, [0], 40, 1e-4)
    assert c[0] == 1
    c, m, _ = F.coincidence_counter([np.array([95, 105])], data, [0], 40, 1e-4)
    assert c[0] == 2                                # two targets, two model spikes
    c, m, _ = F.coincidence_counter([np.array([95, 105])], np.array([-BIG, 100, 5000, BIG]), [0], 40, 1e-4)
    assert c[0] == 1                                # one target spike counts once
    c, m, _ = F.coincidence_counter([np.array([100, 1000])], data, [0], 40, 1e-4, onset=0.05)
    assert c[0] == 1 and m[0] == 1


def test_poisson_rate_function():
    """A rate function of t draws the same spikes as the constant rate it
    returns, and none while it returns 0 (directcontrol.py:436-445)."""
    from oracle import brian1 as O
    dt, N = 1e-4, 50
    a = O.SpikeSource(N, poisson=(np.full(N, 300.0), 9, dt))
    b = O.SpikeSource(N, poisson=(lambda t: 300.0, 9, dt))
    c = O.SpikeSource(N, poisson=(lambda t: 300.0 if t >= 0.01 else 0.0, 9, dt))
    for n in range(300):
        sa, sb, sc = a.update(n * dt, n), b.update(n * dt, n), c.update(n * dt, n)
        np.testing.assert_array_equal(sa, sb)
        if n < 100:
            assert len(sc) == 0
        else:
            np.testing.assert_array_equal(sa, sc)
    assert sum(len(x) for x in a.history) > 200
