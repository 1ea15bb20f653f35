"""Pins of the oracle's random streams and transforms (CPU only).

Philox4x32-10: Random123 published known answers and rocRAND's engine
(tests/golden/philox_rocrand.json, made by make_rocrand_vectors.sh).
u01: closed form.  log/sqrt/cos: libm in float64 within stated ulp bounds
over exhaustive or dense sweeps.  normal: moments and Kolmogorov-Smirnov.
"""
import json
import os

import numpy as np
import pytest
from scipy import stats

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# Random123 kat_vectors, philox4x32 with 10 rounds (Salmon et al. 2011).
R123 = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def ulp_err(approx32, exact64):
    """|approx - exact| in units of the binary32 spacing at `exact`."""
    e32 = exact64.astype(np.float32)
    spacing = np.spacing(np.abs(e32)).astype(np.float64)
    return np.abs(approx32.astype(np.float64) - exact64) / spacing


@pytest.mark.parametrize("ctr,key,out", R123)
def test_philox_random123_kat(oracle, ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out


def test_philox_rocrand_engine(oracle):
    cases = json.load(open(os.path.join(GOLDEN, "philox_rocrand.json")))
    assert len(cases) >= 7
    for c in cases:
        assert list(oracle.philox(c["ctr"], c["key"])) == c["out"], c


def test_philox_many_matches_single(oracle):
    blk = oracle.philox_many(0x123456789ABCDEF0, 10, 77, 4, 5)
    for k in range(5):
        assert list(blk[k]) == list(oracle.philox([10 + k, 77, 4, 0], [0x9ABCDEF0, 0x12345678]))


def test_u01_closed_form(oracle):
    xs = [0, 1, 511, 512, 0x7FFFFFFF, 0x80000000, 0xFFFFFE00, 0xFFFFFFFF, 12345678]
    for x in xs:
        exact = ((x >> 9) * 2 + 1) / 2.0**24
        got = oracle.lib().orc_u01(x)
        assert got == exact and 0.0 < got < 1.0


def test_logf_exhaustive_over_u01(oracle):
    # every value u01 can produce: (2a+1) * 2^-24, a < 2^23
    a = np.arange(1 << 23, dtype=np.float64)
    x = ((2 * a + 1) / 2.0**24).astype(np.float32)
    got = oracle.logf(x)
    # measured maximum 2.42 ulp (near x = 0.97, where (m-1)/(m+1) rounds twice)
    assert ulp_err(got, np.log(x.astype(np.float64))).max() <= 3.0


def test_logf_general(oracle):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-80, 80, 1 << 20)).astype(np.float32)
    x = x[np.isfinite(x) & (x > np.finfo(np.float32).tiny)]
    got = oracle.logf(x)
    exact = np.log(x.astype(np.float64))
    # relative error where log is small (x near 1) needs the absolute form
    err = np.abs(got.astype(np.float64) - exact)
    assert (err <= np.maximum(2 * np.spacing(np.abs(exact.astype(np.float32))), 1e-7)).all()


def test_sqrtf(oracle):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(1e-7, 40, 1 << 20), np.exp(rng.uniform(-80, 80, 1 << 18))]).astype(np.float32)
    x = x[np.isfinite(x) & (x > np.finfo(np.float32).tiny)]
    got = oracle.sqrtf(x)
    assert ulp_err(got, np.sqrt(x.astype(np.float64))).max() <= 1.0


def test_cos2pi_exhaustive(oracle):
    k = np.arange(1 << 24, dtype=np.uint32)
    got = oracle.cos2pi(k << np.uint32(8))
    exact = np.cos(2 * np.pi * (k.astype(np.float64) + 0.5) / 2.0**24)
    assert np.abs(got - exact).max() < 2e-7
    # low 8 bits are ignored
    assert np.array_equal(oracle.cos2pi(k[:4096] << np.uint32(8) | np.uint32(0xFF)), got[:4096])


def test_normal_distribution(oracle):
    w = oracle.philox_many(99, 0, 0, 9, 1 << 21)
    z = oracle.normal(w[:, 0], w[:, 1]).astype(np.float64)
    assert abs(z.mean()) < 4 / np.sqrt(z.size)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / z.size)
    assert abs(stats.skew(z)) < 0.01
    assert abs(stats.kurtosis(z)) < 0.02
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert np.abs(z).max() <= np.sqrt(-2 * np.log(2.0**-24)) + 1e-5


def test_normal_box_muller_form(oracle):
    # z = sqrt(-2 ln u1) * cos(2 pi u2): compare with the float64 formula
    w = oracle.philox_many(5, 0, 0, 9, 1 << 18)
    z = oracle.normal(w[:, 0], w[:, 1]).astype(np.float64)
    u1 = ((w[:, 0] >> 9).astype(np.float64) * 2 + 1) / 2.0**24
    u2 = ((w[:, 1] >> 8).astype(np.float64) + 0.5) / 2.0**24
    exact = np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
    assert np.abs(z - exact).max() < 4e-6


def test_thalamic_stream_independence(oracle):
    a = oracle.thalamic(1, 0, 1 << 16).astype(np.float64)
    b = oracle.thalamic(1, 1, 1 << 16).astype(np.float64)
    c = oracle.thalamic(2, 0, 1 << 16).astype(np.float64)
    for x, y in ((a, b), (a, c), (a[:-1], a[1:])):
        assert abs(np.corrcoef(x, y)[0, 1]) < 5 / np.sqrt(x.size)
