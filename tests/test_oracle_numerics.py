"""Pins the oracle's shared float32 exp and counter RNG (CPU)."""

import numpy as np

from oracle import numerics


def _ulp_err(a, ref):
    a = np.asarray(a, np.float64)
    spacing = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(a - ref) / spacing


def test_expf32_against_float64_exp():
    x = np.linspace(-87.0, 88.0, 400_001).astype(np.float32)
    got = numerics.expf32(x)
    ref = np.exp(x.astype(np.float64))
    assert _ulp_err(got, ref).max() <= 3.0
    assert numerics.expf32(np.float32(0.0)) == 1.0
    assert numerics.expf32(np.float32(-100.0)) == 0.0
    assert np.isinf(numerics.expf32(np.float32(100.0)))
    assert np.isnan(numerics.expf32(np.float32(np.nan)))


def test_expf32_membership_range_dense():
    """The range used by memberships, exp(-t) for t in [0, 20], at fine resolution."""
    x = -np.random.default_rng(0).uniform(0, 20, 1_000_000).astype(np.float32)
    assert _ulp_err(numerics.expf32(x), np.exp(x.astype(np.float64))).max() <= 3.0


def _splitmix_final_py(x):
    m = (1 << 64) - 1
    x &= m
    x ^= x >> 30
    x = (x * 0xBF58476D1CE4E5B9) & m
    x ^= x >> 27
    x = (x * 0x94D049BB133111EB) & m
    x ^= x >> 31
    return x


def test_mix64_matches_python_integers():
    vals = [0, 1, 2, 12345, 2**63 + 17, 2**64 - 1]
    got = numerics.mix64(np.array(vals, dtype=np.uint64))
    assert [int(v) for v in got] == [_splitmix_final_py(v) for v in vals]


def test_coin_stream_uniform_and_keyed():
    key = numerics.stream_key(0, 0)
    u = numerics.coin_u32(key, 0, np.arange(200_000, dtype=np.uint64)).astype(np.float64) / 2**32
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    heads = u < 0.9
    assert abs(heads.mean() - 0.9) < 0.005
    u2 = numerics.coin_u32(key, 1, np.arange(200_000, dtype=np.uint64))
    assert np.mean(u2 == (u * 2**32).astype(np.uint64)) < 1e-3
    k2 = numerics.stream_key(0, 1)
    assert k2 != key
    assert int(numerics.coin_threshold(0.9)) == int(np.floor(0.9 * 2**32))
