"""The FP8 screen's float -> E4M3 rounding (gm_debug_e4m3 on the host; the
device's is compared in test_gpu_fp8.py) against the OCP E4M3 definition:
the nearest finite code, ties to an even mantissa, saturating at 448.  The
reference values come from the format's definition (sign, 4 exponent bits
of bias 7, 3 mantissa bits, subnormals m 2^-9), not from the converter."""

import numpy as np

from gpumap_amd import capi


def _finite_codes():
    codes = np.arange(0x7f, dtype=np.int64)  # 0x7f is NaN
    e, m = codes >> 3, codes & 7
    vals = np.where(e == 0, m * 2.0 ** -9, (1 + m / 8.0) * 2.0 ** (e - 7))
    return codes, vals


def _reference(x):
    codes, vals = _finite_codes()
    a = np.abs(x.astype(np.float64))
    d = np.abs(a[:, None] - vals[None, :])
    best = d.min(axis=1, keepdims=True)
    tie = d == best
    # among the nearest, the even code (even mantissa)
    pick = np.where(tie & (codes[None, :] % 2 == 0), codes[None, :], 1 << 30)
    pick = np.where(tie.sum(1, keepdims=True) > 1, pick,
                    np.where(tie, codes[None, :], 1 << 30))
    c = pick.min(axis=1)
    # at and above 448 the nearest finite code is 448's (the distances above
    # stop resolving it for huge x)
    c = np.where(a >= 448.0, 0x7e, c)
    return (c | np.where(np.signbit(x), 0x80, 0)).astype(np.uint8)


def test_codes_midpoints_and_random_values():
    _, vals = _finite_codes()
    v = vals.astype(np.float32)
    mids = ((vals[:-1] + vals[1:]) / 2).astype(np.float32)  # exact ties
    near = np.concatenate([np.nextafter(mids, np.float32(0)),
                           np.nextafter(mids, np.float32(np.inf))])
    g = np.random.default_rng(46)
    rnd = np.exp(g.uniform(np.log(2.0 ** -14), np.log(600.0), 20000))
    big = np.array([448, 449, 460, 463.99, 464, 500, 1e6, 3.4e38], np.float64)
    x = np.concatenate([v, mids, near, rnd, big,
                        [0.0, 2.0 ** -10, 2.0 ** -11, 1e-30]]).astype(np.float32)
    x = np.concatenate([x, -x])
    got = capi.e4m3(x)
    want = _reference(x)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(want[i]))
                           for i in bad[:10]]
    # every finite code is its own value's rounding
    assert np.array_equal(capi.e4m3(v), np.arange(0x7f, dtype=np.uint8))


def test_rounding_error_bounds_used_by_the_screen():
    # |x - e4m3(x)| <= 2^-4 |x| in the normal range and <= 2^-10 below 2^-6,
    # the two constants of the FP8 screen's error bound
    _, vals = _finite_codes()
    g = np.random.default_rng(47)
    x = np.exp(g.uniform(np.log(2.0 ** -16), np.log(448.0), 50000))
    x = x.astype(np.float32)
    c = capi.e4m3(x)
    e, m = (c >> 3) & 15, c & 7
    dec = np.where(e == 0, m * 2.0 ** -9, (1 + m / 8.0) * 2.0 ** (e - 7.0))
    err = np.abs(dec - x.astype(np.float64))
    assert (err <= 2.0 ** -4 * np.abs(x) + 2.0 ** -10 + 1e-30).all()
    normal = x >= 2.0 ** -6
    assert (err[normal] <= 2.0 ** -4 * x[normal]).all()
    assert (err[~normal] <= 2.0 ** -10).all()
