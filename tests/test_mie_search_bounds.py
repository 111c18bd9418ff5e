"""Mie sampling's search bounds: the kernels narrow MiePhase::Sample's search
(KernelSampler.cpp:258-285, a binary search for the first key >= u1 over
i = 0..254) to bounds per u1 bucket: the answers at the bucket's smallest and
largest uniform, widened by one each way. This checks, over every uniform the
generator can return (odd multiples of 2^-24) and with the kernels' float32
arithmetic, that the bounded search returns the full search's answer."""
import numpy as np

from oracle import pyoracle as O

BUCKETS = 1024


def _search(cdf, u, lo, hi):
    """First i in [lo, hi) with key_i(u) >= u, else hi (the kernels' mie_search)."""
    ub = np.float32(256.0) * u
    lo = lo.astype(np.int64)
    hi = hi.astype(np.int64)
    while True:
        act = lo < hi
        if not act.any():
            return lo
        mid = (lo + hi) >> 1
        midc = np.minimum(mid, 254)
        b = np.maximum(ub, midc.astype(np.float32))
        c = np.minimum(np.floor(b), np.float32(255.0))
        t = (b - c).astype(np.float32)
        c0 = cdf[midc]
        c1 = cdf[midc + 1]
        key = (t * (c1 - c0)).astype(np.float32) + c0
        below = key < u
        lo = np.where(act & below, mid + 1, lo)
        hi = np.where(act & ~below, mid, hi)


def test_bucket_bounds_keep_every_answer():
    cdf = np.asarray(O.MIE_CDF[:, 0], dtype=np.float32)
    m = np.arange(1, 1 << 24, 2, dtype=np.int64)                   # odd multiples of 2^-24
    u = (m.astype(np.float64) * 2.0 ** -24).astype(np.float32)
    assert np.array_equal(u.astype(np.float64) * 2.0 ** 24, m)       # exactly representable
    full = _search(cdf, u, np.zeros(m.size), np.full(m.size, 255))
    # The answer mostly grows with u1; where it steps back, by one only.
    assert np.diff(full).min() >= -1
    per = (1 << 24) // BUCKETS
    b = np.arange(BUCKETS)
    z, e = np.zeros(BUCKETS), np.full(BUCKETS, 255)
    lo_b = _search(cdf, ((b * per + 1) * 2.0 ** -24).astype(np.float32), z, e)
    hi_b = _search(cdf, (((b + 1) * per - 1) * 2.0 ** -24).astype(np.float32), z, e)
    lo_b = np.maximum(lo_b - 1, 0)
    hi_b = np.minimum(hi_b + 1, 255)
    bucket = (u * np.float32(BUCKETS)).astype(np.int64)
    assert np.array_equal(bucket, m // per)
    bounded = _search(cdf, u, lo_b[bucket], hi_b[bucket])
    assert np.array_equal(bounded, full)
    assert np.mean(hi_b - lo_b) < 6   # a few steps instead of eight
