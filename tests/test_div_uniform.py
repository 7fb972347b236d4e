"""mx::div_uniform (device/mapx_dev.hiph): a floating division by a divisor
that is uniform over the work-group becomes a multiplication by the exact
reciprocal when the divisor is a power of two whose reciprocal is normal.

The claim checked here is the one the rewrite rests on: for every divisor
the bit test accepts, x * reciprocal equals the IEEE quotient x / d bit for
bit, over normal, subnormal, zero, infinite and NaN dividends; and the test
accepts exactly the powers of two in that range. Both follow the bit
operations of the device code on IEEE binary32 / binary64 in numpy.
"""
import numpy as np

SPEC = {np.float32: (np.uint32, 23, 0xff, 254), np.float64: (np.uint64, 52, 0x7ff, 2046)}


def recip_of(d, ft):
    ut, mb, emask, top = SPEC[ft]
    b = int(np.array(d, ft).view(ut))
    e = (b >> mb) & emask
    if (b & ((1 << mb) - 1)) == 0 and 1 <= e <= top - 1:
        sign = b & (1 << (mb + 1 + (8 if ft is np.float32 else 11) - 1))
        return np.array(sign | ((top - e) << mb), ut).view(ft)
    return None


def dividends(ft, rng):
    fi = np.finfo(ft)
    x = np.concatenate([rng.standard_normal(2000).astype(ft) * ft(1e3),
                        (rng.random(500) * fi.tiny).astype(ft),  # subnormal
                        np.array([0.0, -0.0, fi.max, -fi.max, fi.tiny, fi.smallest_subnormal,
                                  np.inf, -np.inf, np.nan], ft),
                        (rng.random(500).astype(ft) * fi.max).astype(ft)])
    return x


def test_power_of_two_divisors_divide_exactly():
    rng = np.random.default_rng(11)
    for ft in (np.float32, np.float64):
        ut, mb, emask, top = SPEC[ft]
        x = dividends(ft, rng)
        accepted = 0
        for e in range(0, emask + 1):
            for sign in (0, 1):
                bits = (sign << (mb + (8 if ft is np.float32 else 11))) | (e << mb)
                d = np.array(bits, ut).view(ft)
                r = recip_of(d, ft)
                normal_recip = 1 <= e <= top - 1
                assert (r is not None) == normal_recip, (ft, e)
                if r is None:
                    continue
                accepted += 1
                assert float(r) * float(d) == 1.0
                with np.errstate(all="ignore"):
                    q = x / d
                    p = x * r
                assert np.array_equal(q.view(ut), p.view(ut)) or np.array_equal(q, p, equal_nan=True), (ft, e)
        assert accepted == 2 * (top - 1)


def test_other_divisors_are_not_rewritten():
    for ft in (np.float32, np.float64):
        for d in (3.0, 0.1, 10.0, 8.000001, -7.0, 0.0, np.inf, np.nan, np.finfo(ft).tiny / 4):
            assert recip_of(ft(d), ft) is None, (ft, d)
        assert recip_of(ft(8.0), ft) == ft(0.125) and recip_of(ft(-0.5), ft) == ft(-2.0)
