"""Pins of the oracle's SGD (oracle/oracle.cpp sgd_epochs, restating
embedding_kernel.cuh:44-75,325-589,726-798) against what fixes each part of it
independently of the code:

- the gradient coefficients against numerical derivatives of the UMAP loss
  terms (attraction log(1 + a s^b), repulsion -gamma log(1 - 1/(1 + a s^b)),
  with the reference's 0.001 regulariser), so a dropped factor, a wrong sign
  or a missing gamma fails;
- the sampling schedule against its closed forms (the m-th positive sample of
  an edge of relative weight r at epoch ceil(m / r); nsr * r negative draws
  per epoch, cumulative floor(t nsr r) - 1 after a sample at epoch t);
- the negative rows against the published Philox4x32-10 (itself pinned by
  Random123's known answers) and the multiply-high index map;
- the learning rate against alpha0 (1 - (t - 1) / E);
- the rounding grid against its definition (the smallest power of two above
  the largest sum a row can receive), and the steps' composition into the
  output through that grid.
CPU only."""

import math

import numpy as np
import pytest

import oracle

A, B = 1.577, 0.895
TAG_NEG = 4  # counter word 3 of the negative-sample streams


def _attr_coef_fd(s, a=A, b=B):
    """-2 dL/ds of L(s) = log(1 + a s^b), by central differences (FP64)."""
    f = lambda v: math.log1p(a * v ** b)  # noqa: E731
    h = s * 1e-6
    return -2.0 * (f(s + h) - f(s - h)) / (2 * h)


def _rep_coef_fd(s, a=A, b=B, gamma=1.0):
    """-2 dL/ds of L(s) = -gamma log(1 - 1 / (1 + a s^b)), central
    differences (FP64)."""
    f = lambda v: -gamma * math.log(1.0 - 1.0 / (1.0 + a * v ** b))  # noqa
    h = s * 1e-6
    return -2.0 * (f(s + h) - f(s - h)) / (2 * h)


def _neg_rows(seed, epoch, edge, count, n):
    """Rows of an edge's negative samples in an epoch: Philox4x32-10 with key
    seed + epoch and counter (edge, edge >> 32, block, TAG_NEG), the draws in
    order, index = high word of draw * n."""
    key64 = seed + epoch
    key = (key64 & 0xFFFFFFFF, key64 >> 32)
    out, blk = [], 0
    while len(out) < count:
        out += [int(v) for v in oracle.philox4x32(
            (edge & 0xFFFFFFFF, edge >> 32, blk, TAG_NEG), key)]
        blk += 1
    return [(w * n) >> 32 for w in out[:count]]


def _grid_add(r, v):
    """v truncated to the grid of the rounding factor r, then added (float32):
    (r + v) - r."""
    r32 = np.float32(r)
    return (r32 + np.float32(v)) - r32


def _frac_margin(x):
    return abs(x - round(x))


# relative weights chosen so that m / r and t nsr r stay clear of integers
# (or are exact), so float rounding cannot move a sample by an epoch
W_REL = [1.0, 0.3713, 0.6171, 0.2459, 0.8133]
E, NSR = 40, 4


def _schedule_case(gamma=1.0, seed=11):
    g = np.random.default_rng(21)
    n = 12
    emb = g.uniform(-3, 3, size=(n, 2)).astype(np.float32)
    head = np.array([0, 2, 4, 1, 7], np.int32)
    tail = np.array([1, 3, 5, 6, 0], np.int32)
    w = np.array(W_REL, np.float32) * np.float32(0.8)
    out, ev, r = oracle.sgd_trace(emb, head, tail, w, E, NSR, A, B,
                                  alpha0=1.0, gamma=gamma, seed=seed)
    return emb, head, tail, w, out, ev, r, n, seed, gamma


def test_weights_keep_clear_of_float_ties():
    for rr in W_REL:
        eps = float(np.float32(1.0) / np.float32(rr))
        for m in range(1, E + 1):
            x = m * eps
            if x > E:
                break
            assert x == round(x) or _frac_margin(x) > 1e-3, (rr, m)
        for t in range(1, E):
            x = t * NSR * rr
            assert x == round(x) or _frac_margin(x) > 1e-3, (rr, t)


@pytest.mark.parametrize("gamma", [1.0, 0.5])
def test_sgd_steps_follow_the_loss_schedule_and_draws(gamma):
    emb, head, tail, w, out, ev, rgrid, n, seed, gamma = _schedule_case(gamma)
    mw = float(w.max())
    pos = ev[ev["kind"] == oracle.SGD_POSITIVE]
    negs = ev[ev["kind"] != oracle.SGD_POSITIVE]
    for e in range(len(head)):
        rr = float(w[e]) / mw
        eps = float(np.float32(1.0) / np.float32(rr))
        # positive samples: the m-th at epoch ceil(m eps)
        want = [math.ceil(m * eps - 1e-9) for m in range(1, E + 1)
                if math.ceil(m * eps - 1e-9) < E]
        got = pos[pos["edge"] == e]["epoch"].tolist()
        assert got == want, (e, got, want)
        # negative draws: cumulative floor(t nsr r) - 1 after the sample at t
        cum = 0
        for t in got:
            cnt = int(((negs["edge"] == e) & (negs["epoch"] == t)).sum())
            cum += cnt
            assert cum == math.floor(t * NSR * rr + 1e-9) - 1, (e, t, cum)
            # their rows: the published generator's draws, multiply-high map
            rows = _neg_rows(seed, t, e, cnt, n)
            sel = negs[(negs["edge"] == e) & (negs["epoch"] == t)]
            assert sel["other"].tolist() == rows
            skipped = sel["kind"] == oracle.SGD_NEGATIVE_SKIPPED
            assert (np.array(rows)[skipped] == head[e]).all()
            assert (np.array(rows)[~skipped] != head[e]).all()
    # the learning rate decays linearly, one step behind the epoch
    for x in ev:
        want = 1.0 if x["epoch"] == 0 else 1.0 - (x["epoch"] - 1) / E
        assert x["alpha"] == pytest.approx(want, rel=1e-6)
    # gradient coefficients: numerical derivatives of the loss terms
    for x in pos:
        y, o = x["y"][:2].astype(np.float64), x["o"][:2].astype(np.float64)
        s = float(((y - o) ** 2).sum())
        assert x["d2"] == pytest.approx(s, rel=1e-5)
        assert x["coef"] == pytest.approx(_attr_coef_fd(s), rel=2e-4)
    live = negs[negs["kind"] == oracle.SGD_NEGATIVE]
    assert len(live) > 200
    for x in live:
        y, o = x["y"][:2].astype(np.float64), x["o"][:2].astype(np.float64)
        s = float(((y - o) ** 2).sum())
        assert x["d2"] == pytest.approx(s, rel=1e-5)
        # the reference regularises the 1/s factor as 1/(0.001 + s)
        coef = float(x["coef"]) * (0.001 + s) / s
        assert coef == pytest.approx(_rep_coef_fd(s, gamma=gamma), rel=2e-4)
    # each step: clip(coef (y - o), +-4) alpha per component
    for x in ev[ev["kind"] != oracle.SGD_NEGATIVE_SKIPPED]:
        for d in range(2):
            v = np.float32(x["coef"]) * (x["y"][d] - x["o"][d])
            want = np.float32(np.clip(v, -4, 4)) * np.float32(x["alpha"])
            assert x["g"][d] == pytest.approx(want, rel=2e-6, abs=1e-12)


def test_sgd_steps_compose_into_the_output_through_the_rounding_grid():
    emb, head, tail, w, out, ev, rgrid, n, seed, _ = _schedule_case()
    # grid: smallest power of two above max_deg 4 alpha / (1 - 2 max_deg eps)
    maxdeg = np.bincount(head, minlength=n).max()
    delta = maxdeg * 4.0 / (1 - 2 * maxdeg * 2.0 ** -23)
    assert rgrid == 2.0 ** (math.floor(math.log2(delta)) + 1)
    # replay the recorded steps: each sample reads the rows as the samples
    # before it left them; the tail takes the truncated negative of the
    # attraction at once, the head the truncated sum of its steps at the end
    cur = emb.copy()
    i = 0
    while i < len(ev):
        x = ev[i]
        assert x["kind"] == oracle.SGD_POSITIVE
        e = x["edge"]
        h, t = head[e], tail[e]
        np.testing.assert_array_equal(x["y"][:2], cur[h])
        np.testing.assert_array_equal(x["o"][:2], cur[t])
        y = cur[h].copy()
        gsum = np.zeros(2, np.float32)
        for d in range(2):
            y[d] = np.float32(y[d] + x["g"][d])
            gsum[d] = x["g"][d]
            cur[t, d] = cur[t, d] + _grid_add(rgrid, -x["g"][d])
        i += 1
        while i < len(ev) and ev[i]["kind"] != oracle.SGD_POSITIVE:
            q = ev[i]
            assert q["edge"] == e
            if q["kind"] == oracle.SGD_NEGATIVE:
                np.testing.assert_array_equal(q["y"][:2], y)
                np.testing.assert_array_equal(q["o"][:2], cur[q["other"]])
                for d in range(2):
                    y[d] = np.float32(y[d] + q["g"][d])
                    gsum[d] = np.float32(gsum[d] + q["g"][d])
            i += 1
        for d in range(2):
            cur[h, d] = cur[h, d] + _grid_add(rgrid, gsum[d])
    np.testing.assert_array_equal(cur, out)


@pytest.mark.parametrize("maxdeg", [1, 2, 3, 5, 8, 100])
@pytest.mark.parametrize("alpha", [1.0, 0.37])
def test_rounding_factor_is_the_smallest_power_of_two_above_the_bound(maxdeg,
                                                                      alpha):
    n = 400
    head = np.concatenate([np.zeros(maxdeg, np.int32),
                           np.arange(1, 1 + min(maxdeg, 50), dtype=np.int32)])
    r = oracle.rounding_factor(head, n, alpha)
    delta = maxdeg * 4.0 * alpha / (1 - 2 * maxdeg * 2.0 ** -23)
    assert r == 2.0 ** (math.floor(math.log2(delta)) + 1)
    assert r > delta >= r / 2
    # why: contributions truncated through r (multiples of r 2^-24, the
    # spacing of floats just below r), each at most 4 alpha, sum exactly in
    # float32 in any order when at most maxdeg of them meet
    g = np.random.default_rng(maxdeg)
    v = [_grid_add(r, x) for x in g.uniform(-4 * alpha, 4 * alpha, maxdeg)]
    exact = sum(float(x) for x in v)
    acc = np.float32(0.0)
    for x in v[::-1]:
        acc = np.float32(acc + x)
    assert float(acc) == exact
    for x in v:
        assert (float(x) / (r * 2.0 ** -24)).is_integer()


def test_clipped_and_coincident_steps():
    # two rows 0.01 apart: the repulsion from the tail is far above 4 and is
    # clipped; then rows at one point: the attraction vanishes (d2 = 0) and a
    # negative sample at the head's own position moves it by 4 alpha
    for emb, kind in ((np.array([[0.0, 0.0], [0.01, 0.0]], np.float32), "clip"),
                      (np.array([[0.5, 0.5], [0.5, 0.5]], np.float32), "zero")):
        out, ev, r = oracle.sgd_trace(emb, [0], [1], np.ones(1, np.float32),
                                      6, 4, A, B, seed=0)
        live = ev[ev["kind"] == oracle.SGD_NEGATIVE]
        assert len(live)
        if kind == "clip":
            x = live[0]
            v = float(x["coef"]) * float(x["y"][0] - x["o"][0])
            assert abs(v) > 4
            assert abs(x["g"][0]) == pytest.approx(4 * x["alpha"], rel=1e-7)
        else:
            p = ev[ev["kind"] == oracle.SGD_POSITIVE][0]
            assert p["d2"] == 0 and p["coef"] == 0 and (p["g"][:2] == 0).all()
            x = live[0]
            assert x["d2"] == 0.0
            np.testing.assert_array_equal(x["g"][:2], [4 * x["alpha"]] * 2)
