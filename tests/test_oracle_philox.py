"""Philox4x32-10 known-answer vectors (Random123, Salmon et al. SC'11) and
the statistics of the samplers drawn from it."""
import numpy as np
import pytest
from scipy import stats

from oracle import philox as op

KAT = [  # (counter, key) -> output, Random123 kat_vectors for philox4x32_10
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox4x32_10_known_answers(ctr, key, out):
    got = op.philox4x32(np.array(ctr, dtype=np.uint32), np.array(key, dtype=np.uint32))
    assert tuple(int(v) for v in got) == out


def test_uniform_is_open_interval_and_uniform():
    w = op.block(np.arange(200000), 0, 3, 1, 12345)[..., 0]
    u = op.uniform(w)
    assert u.min() > 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3


def test_normal_draws_are_standard_normal():
    z = op.normal(np.arange(200000), 5, 1, 99)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3


@pytest.mark.parametrize("lam", [0.3, 4.0, 25.0, 900.0])
def test_poisson_draws_follow_the_poisson_law(lam):
    n = 20000
    k = op.poisson(np.full(n, lam), 0, 7, 0, 2024)
    assert np.all(k == np.floor(k)) and k.min() >= 0
    se = np.sqrt(lam / n)
    assert abs(k.mean() - lam) < 5 * se
    assert abs(k.var() / lam - 1.0) < 0.06
    # cumulative frequencies at a few quantiles of the exact law
    for q in (0.1, 0.5, 0.9):
        kq = stats.poisson.ppf(q, lam)
        p = stats.poisson.cdf(kq, lam)
        assert abs(np.mean(k <= kq) - p) < 5 * np.sqrt(p * (1 - p) / n)


def test_streams_are_separated_by_counter_fields():
    a = op.block(np.arange(64), 0, 0, 0, 1)
    for other in (op.block(np.arange(64), 1, 0, 0, 1), op.block(np.arange(64), 0, 1, 0, 1),
                  op.block(np.arange(64), 0, 0, 1, 1), op.block(np.arange(64), 0, 0, 0, 2)):
        assert not np.any(np.all(a == other, axis=-1))


def _device_ptrs_thresholds():
    """The integer thresholds the device sampler (rng_poisson_f in
    shwfs_device.hpp) uses in place of the oracle's float64 tests
    us >= 0.07 and us < 0.013, read from the source."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "compass-saxoplus_amd", "csrc",
                            "shwfs_device.hpp")).read()
    hi = re.search(r"if \(m >= (\d+)u\) \{\s*// us >= 0\.07", src)
    lo = re.search(r"\(m <= (\d+)u && wv > m\)\) continue;\s*// us < 0\.013 && V > us", src)
    assert hi and lo, "threshold lines not found in shwfs_device.hpp"
    return int(hi.group(1)), int(lo.group(1))


def _oracle_us(wu):
    # the oracle's arithmetic (philox._poisson_one): U = uniform - 0.5, us = 0.5 - |U|
    U = op.uniform(np.asarray(wu, dtype=np.uint64)) - 0.5
    return 0.5 - np.abs(U)


def test_ptrs_integer_tests_equal_the_oracle_float_tests():
    """us = (min(wu, ~wu) + 1/2) 2^-32 exactly, so the oracle's tests
    us >= 0.07, us < 0.013 and V > us are integer tests on the Philox words:
    checked for random words and for every word within 4 of each threshold
    (both halves of the word range), against the oracle's float64 values."""
    hi, lo = _device_ptrs_thresholds()
    rng = np.random.default_rng(5)
    words = rng.integers(0, 2 ** 32, 200000, dtype=np.uint64)
    near = []
    for t in (hi, lo):
        for d in range(-4, 5):
            m = t + d
            near += [m, 2 ** 32 - 1 - m]  # wu below and above 2^31 with min(wu, ~wu) = m
    words = np.concatenate([words, np.array(near, dtype=np.uint64)])
    us = _oracle_us(words)
    m = np.minimum(words, np.uint64(2 ** 32 - 1) - words)
    np.testing.assert_array_equal(us, (m.astype(np.float64) + 0.5) * 2.0 ** -32)
    np.testing.assert_array_equal(us >= 0.07, m >= hi)
    np.testing.assert_array_equal(us < 0.013, m <= lo)
    wv = rng.integers(0, 2 ** 32, words.size, dtype=np.uint64)
    wv[-len(near):] = m[-len(near):]  # V == us exactly at the threshold words
    wv[:1000] = m[:1000]              # and at some random ones
    wv[1000:2000] = m[1000:2000] + 1  # one step above
    V = op.uniform(wv)
    np.testing.assert_array_equal(V > us, wv > m)
