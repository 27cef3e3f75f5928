"""Shared float32 / integer numerics of the oracle — TEST INFRASTRUCTURE ONLY.

Two definitions are restated identically in ``metal-singlecell_amd/csrc/common.h`` so that
the HIP kernels and this oracle produce bit-identical results:

``expf32``   exp(x) in float32 by Cody-Waite range reduction (x = n ln2 + r, with the
             fdlibm two-part ln2 split) and the degree-7 Taylor polynomial of e^r evaluated
             by Horner's rule, every operation a separately rounded float32 multiply or
             add (no fused multiply-add).  Results below 2^-126 flush to 0.  Error versus
             the true exponential is a few ulp (pinned in tests/test_oracle_numerics.py).

``coin_u32`` a counter-based generator: splitmix64's finaliser (Steele, Lea & Flood 2014)
             applied to (seed, phase, pass, vertex).  Leiden's random half-commit draws one
             u32 per vertex per pass from it; the coin is ``u32 < floor(p * 2^32)``.
"""

from __future__ import annotations

import numpy as np

F = np.float32
LOG2E = F(1.44269504088896341)
LN2_HI = F(6.93145751953125e-01)      # 0x3f317200: 16 significant bits, n*LN2_HI exact
LN2_LO = F(1.42860682030941723e-06)   # 0x35bfbe8e
# 1/j! for j = 7..1, rounded to float32
_TAYLOR = [F(1.0 / 5040.0), F(1.0 / 720.0), F(1.0 / 120.0), F(1.0 / 24.0), F(1.0 / 6.0),
           F(0.5), F(1.0)]
EXP_LO = F(-87.0)          # below: result < 2^-125.5 -> flushed to zero
EXP_HI = F(88.0)


def expf32(x):
    """Elementwise float32 exp with the project's fixed operation sequence."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        xc = np.clip(x, EXP_LO, EXP_HI).astype(np.float32)
        nf = np.rint(xc * LOG2E).astype(np.float32)
        r = (xc - nf * LN2_HI).astype(np.float32)
        r = (r - nf * LN2_LO).astype(np.float32)
        p = np.full_like(r, _TAYLOR[0])
        for c in _TAYLOR[1:]:
            p = (p * r).astype(np.float32)
            p = (p + c).astype(np.float32)
        p = (p * r).astype(np.float32)
        p = (p + F(1.0)).astype(np.float32)
        out = np.ldexp(p, nf.astype(np.int32)).astype(np.float32)
    out = np.where(x < EXP_LO, F(0.0), out)
    out = np.where(x > EXP_HI, F(np.inf), out)
    out = np.where(np.isnan(x), F(np.nan), out)
    return out.astype(np.float32)


M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
PASS_MUL = np.uint64(0xD1B54A32D192ED03)


def mix64(x):
    """splitmix64 finaliser on uint64 (arrays or scalars), wrapping arithmetic."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x ^ (x >> np.uint64(30))
        x = x * np.uint64(0xBF58476D1CE4E5B9)
        x = x ^ (x >> np.uint64(27))
        x = x * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def stream_key(seed, phase):
    """Per-(seed, phase) key; phase 0 = local moving, 1 = refinement."""
    with np.errstate(over="ignore"):
        s = np.uint64(np.int64(seed) & np.int64(0x7FFFFFFFFFFFFFFF))
        return mix64(s * GOLDEN + np.uint64(phase))


def coin_u32(key, pass_idx, v):
    """u32 draw for every vertex in ``v`` at pass ``pass_idx`` of the stream ``key``."""
    with np.errstate(over="ignore"):
        h = mix64(np.uint64(key) + np.uint64(pass_idx + 1) * PASS_MUL)
        u = mix64(h ^ (np.asarray(v, dtype=np.uint64) * GOLDEN))
    return (u >> np.uint64(32)).astype(np.uint64)


def coin_threshold(p):
    """Integer threshold: the coin is heads when ``u32 < coin_threshold(p)``."""
    return np.uint64(min(int(np.floor(float(p) * 4294967296.0)), 4294967296))
