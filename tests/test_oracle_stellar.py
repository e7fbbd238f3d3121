"""The oracle's star formation (nbody_bh.wgsl:197-268) pinned without the
reference: the hash against an independent statement of the published PCG
construction and its permutation property, the IMF sampler against the
closed-form Kroupa inverse CDF (endpoints, branch weight, mean), and the
formation / recycling rules against hand-built particles with closed-form
outcomes."""
import numpy as np
import pytest

from oracle import oracle

M32 = 0xFFFFFFFF


def pcg_published(v):
    """PCG32 (O'Neill 2014): LCG step state*747796405 + 2891336453, then the
    RXS-M-XS output function for 32-bit state (pcg-cpp rxs_m_xs_mixin,
    multiplier 277803737), written from that description."""
    state = (v * 747796405 + 2891336453) & M32
    rshift = (state >> 28) + 4
    word = (((state >> rshift) ^ state) * 277803737) & M32
    return (word >> 22) ^ word


def test_pcg_matches_published_construction_and_is_a_permutation():
    xs = np.arange(0, 1 << 16, dtype=np.uint32) * np.uint32(65537) + np.uint32(12345)
    got = oracle.pcg(xs[:4096])
    want = np.array([pcg_published(int(x)) for x in xs[:4096]], np.uint32)
    assert np.array_equal(got, want)
    # a bijection of u32: no collisions over 2^16 distinct inputs
    assert np.unique(np.array([pcg_published(int(x)) for x in xs])).size == xs.size


def test_rand01_is_uniform_on_unit_interval():
    u = oracle.rand01(np.arange(200000, dtype=np.uint32), 12345)
    assert u.min() >= 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1.0 / 12.0) < 2e-3
    v = oracle.rand01(np.arange(200000, dtype=np.uint32), 12345 + 7)
    assert abs(np.corrcoef(u, v)[0, 1]) < 1e-2  # offset salts give separate streams


def kroupa_branch_weight():
    low = (0.08 ** -0.3 - 0.5 ** -0.3) / 0.3
    hi = 0.5 * (0.5 ** -1.3 - 100.0 ** -1.3) / 1.3
    return low / (low + hi)


def test_imf_sample_is_the_kroupa_inverse_cdf():
    p_low = kroupa_branch_weight()
    lo, hi = 0.5 * p_low, 0.5 * (1 + p_low)  # u1 inside each branch
    assert oracle.imf_sample(lo, 0.0) == pytest.approx(0.08, rel=1e-5)
    assert oracle.imf_sample(lo, 1.0) == pytest.approx(0.5, rel=1e-5)
    assert oracle.imf_sample(hi, 0.0) == pytest.approx(0.5, rel=1e-5)
    assert oracle.imf_sample(hi, 1.0) == pytest.approx(100.0, rel=1e-4)
    u2 = np.linspace(0, 1, 101)
    m_low = [oracle.imf_sample(lo, x) for x in u2]
    m_hi = [oracle.imf_sample(hi, x) for x in u2]
    assert np.all(np.diff(m_low) > 0) and np.all(np.diff(m_hi) > 0)
    # mean mass of the two-segment power law dN/dm ~ m^-1.3 on [0.08,0.5],
    # 0.5 m^-2.3 on [0.5,100] (continuous at 0.5), against quasi-random draws
    rng = np.random.default_rng(1)
    u = rng.random((40000, 2))
    ms = np.array([oracle.imf_sample(a, b) for a, b in u])
    num = (0.5 ** 0.7 - 0.08 ** 0.7) / 0.7 + 0.5 * (0.5 ** -0.3 - 100.0 ** -0.3) / 0.3
    den = (0.08 ** -0.3 - 0.5 ** -0.3) / 0.3 + 0.5 * (0.5 ** -1.3 - 100.0 ** -1.3) / 1.3
    assert ms.mean() == pytest.approx(num / den, rel=0.03)
    assert np.mean(ms < 0.5) == pytest.approx(p_low, abs=0.01)


def _records(pads, vel=(0.0, 0.0, 0.0)):
    n = len(pads)
    r = np.zeros((n, 8), np.float32)
    r[:, 0] = np.arange(n)
    r[:, 3] = 1.0 / n
    r[:, 4:7] = vel
    r[:, 7] = pads
    return r


def test_dense_gas_turns_into_stars_and_thin_gas_stays():
    # density = -pad - 1; threshold 10; sfEff * sfDt * 3 >= 1 makes formation certain
    pads = np.array([-(1 + 50.0), -(1 + 5.0), 0.0, -(1 + 1e4)], np.float32)
    prm = oracle.pack_params(4, time=2.0, aging_rate=25.0, sf_density=10.0, sf_eff=1.0,
                             sf_dt=1.0)
    st = np.zeros((4, 4), np.float32)
    pr, so = oracle.star_formation(_records(pads), st, prm)
    for i in (0, 3):
        assert pr[i, 7] == 0.0
        assert 0.08 <= so[i, 0] <= 100.0
        assert so[i, 1] == np.float32(2.0) * np.float32(25.0)
        assert 0.0 <= so[i, 2] <= 1.0 and so[i, 3] == 0.0
    assert pr[1, 7] == pads[1] and np.all(so[1] == 0)  # below the threshold
    assert pr[2, 7] == 0.0 and np.all(so[2] == 0)      # a star stays a star


def test_massive_stars_recycle_after_their_lifetime():
    # m = 10: tMS = 10000 * 10^-2.5 = 31.62 Myr, tGiant = 0.12 tMS + 15; total + 45
    tms = 10000.0 * 10.0 ** -2.5
    life = tms + (0.12 * tms + 15.0) + 15.0 + 30.0
    now = 4.0 * 25.0  # time 4, agingRate 25 -> 100 Myr
    st = np.array([[10.0, now - life - 1.0, 0.25, 0.0],    # older than its lifetime
                   [10.0, now - life + 1.0, 0.5, 0.0],     # younger
                   [5.0, -1000.0, 0.75, 0.0]], np.float32)  # below 8: never recycled
    prm = oracle.pack_params(3, time=4.0, aging_rate=25.0)
    pr, so = oracle.star_formation(_records(np.zeros(3, np.float32), vel=(1.0, 2.0, 3.0)), st,
                                   prm)
    assert pr[0, 7] == -1.0 and list(so[0]) == [-1.0, 0.0, np.float32(0.25), np.float32(0.7)]
    kick = pr[0, 4:7] - np.array([1.0, 2.0, 3.0], np.float32)
    assert np.all(np.abs(kick) <= 1.0) and np.any(kick != 0)
    assert np.array_equal(pr[1:], _records(np.zeros(3, np.float32), vel=(1.0, 2.0, 3.0))[1:])
    assert np.array_equal(so[1:], st[1:])
