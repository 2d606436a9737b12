"""The euclidean and pearson slab keys' error bounds, checked against an
emulation of the arithmetic on the CPU.

The slab computes x = sum hi_q hi_r + hi_q lo_r + lo_q hi_r (hi = bf16(v),
lo = bf16(v - hi), products exact in FP32, FP32 accumulation) and keys each
pair with minus a lower bound of the squared distance (euc_key in
sparsedist_amd/csrc/k_slab.hip) or an upper bound of the correlation
(pear_key), with constants from dot_bound in sparsedist_amd/csrc/sd_api.cpp.  Here the same rounding steps are emulated
with numpy (bf16 by bit rounding of FP32, the accumulation as a sequential
FP32 cumulative sum in a random order) and the exact dot products come from
math.fsum, so the check does not reuse the formula it tests:

* |x - q.r| <= (e0 + e1 m) sum |q_i r_i| for nonnegative and signed data;
* the euclidean key's lower bound never exceeds the exact squared distance,
  and the pearson key's upper bound is never below the exact correlation;
* a slab that dropped the lo.hi term breaks the first bound (the check has
  teeth for the mistake it guards against);
* the plain BF16 layout (hi only, used for non-negative data) meets its own,
  wider constants.
"""
import math

import numpy as np
import pytest

U23 = 2.0 ** -23
UB = 2.0 ** -8 + U23
K_EUC_ROUND = np.float32(2.0 ** -21)
K_EUC_ABS = np.float32(1e-30)


def bf16(x32):
    """Round FP32 to BF16 (round to nearest even), returned as FP32."""
    b = np.asarray(x32, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def split(v):
    hi = bf16(v.astype(np.float32))
    lo = bf16((v - hi.astype(np.float64)).astype(np.float32))
    return hi, lo


def constants(nrow, terms=3.0):
    n = terms * nrow
    g = 1.0 / (1.0 - n * U23)
    e0 = 3.1 * UB * UB
    e1 = terms * 1.02 * U23 * g
    return e0, e1


def slab_dot(q, r, rng, drop_lo_hi=False):
    qh, ql = split(q)
    rh, rl = split(r)
    parts = [qh.astype(np.float64) * rh, qh.astype(np.float64) * rl]
    if not drop_lo_hi:
        parts.append(ql.astype(np.float64) * rh)
    t = np.concatenate(parts).astype(np.float32)  # products of BF16 pairs are exact
    t = t[t != 0]
    rng.shuffle(t)
    return float(np.cumsum(t, dtype=np.float32)[-1]) if t.size else 0.0


def pair(rng, nrow, nnz, overlap, signed):
    rows_q = rng.choice(nrow, nnz, replace=False)
    keep = rows_q[: int(overlap * nnz)]
    rest = np.setdiff1d(np.arange(nrow), rows_q)
    rows_r = np.concatenate([keep, rng.choice(rest, nnz - keep.size, replace=False)])
    q = np.zeros(nrow)
    r = np.zeros(nrow)
    # log1p-like magnitudes, with values just past BF16 rounding ties mixed in
    vq = np.log1p(rng.gamma(0.5, 4.0, nnz) * 50) + 0.01
    vr = np.log1p(rng.gamma(0.5, 4.0, nnz) * 50) + 0.01
    tie = rng.random(nnz) < 0.3
    vq[tie] = (1.0 + 2.0 ** -8 + 2.0 ** -20) * 2.0 ** rng.integers(-3, 3, tie.sum())
    if signed:
        vq *= rng.choice([-1.0, 1.0], nnz)
        vr *= rng.choice([-1.0, 1.0], nnz)
    q[rows_q] = vq
    r[rows_r] = vr
    return q, r


def key_lower_bound(qa, ra, x, m, e0, e1, nonneg):
    """euc_weights + euc_key's arithmetic in FP32, returning the lower bound
    (minus the key)."""
    f = np.float32
    qa, ra, x = f(qa), f(ra), f(x)
    e = f(f(e0) + f(f(e1) * f(m)))
    w1 = f(f(2.0) * e) if nonneg else f(0.0)
    w2 = K_EUC_ROUND if nonneg else f(e + K_EUC_ROUND)
    n2 = f(qa + ra)
    s = f(n2 - f(f(2.0) * x))
    err = f(f(f(w1 * f(abs(x))) + f(w2 * n2)) + K_EUC_ABS)
    return -float(f(err - s))


@pytest.mark.parametrize("signed", [False, True])
def test_dot_bound_and_lower_bound(signed):
    rng = np.random.default_rng(7 + signed)
    nrow = 4000
    e0, e1 = constants(nrow)
    emax = e0 + e1 * nrow
    fold = 1.0 / (1.0 - emax)
    for trial in range(60):
        nnz = int(rng.integers(50, 1500))
        q, r = pair(rng, nrow, nnz, rng.random(), signed)
        x = slab_dot(q, r, rng)
        exact = math.fsum(q * r)
        s_abs = math.fsum(np.abs(q * r))
        assert abs(x - exact) <= (e0 + e1 * nnz) * s_abs + 1e-300, trial
        d2 = math.fsum((q - r) ** 2)
        lb = key_lower_bound(math.fsum(q * q), math.fsum(r * r), x, nnz, e0 * fold, e1 * fold,
                             not signed)
        assert lb <= d2, (trial, lb, d2)


def test_missing_cross_term_is_caught():
    rng = np.random.default_rng(11)
    nrow = 4000
    e0, e1 = constants(nrow)
    broken = 0
    for trial in range(20):
        q, r = pair(rng, nrow, 800, 0.9, False)
        x = slab_dot(q, r, rng, drop_lo_hi=True)
        exact = math.fsum(q * r)
        s_abs = math.fsum(np.abs(q * r))
        broken += abs(x - exact) > (e0 + e1 * 800) * s_abs
    assert broken >= 15


K_PEAR_ROUND = np.float32(2.0 ** -20)


def pear_upper_bound(qa, qc, ra, rc, x, m, e0, e1, nonneg):
    """pear_weights + pear_key's arithmetic in FP32."""
    f = np.float32
    qa, qc, ra, rc, x = f(qa), f(qc), f(ra), f(rc), f(x)
    ew = f(f(f(e0) + f(f(e1) * f(m))) + K_PEAR_ROUND)
    gq = f(np.sqrt(f(f(1.0) + f(f(qa * qc) * f(qa * qc)))))
    w1 = ew if nonneg else f(0.0)
    w2 = f(0.0) if nonneg else f(ew * gq)
    t = f(qa * ra)
    sc = f(qc * rc)
    r = f(f(x - t) * sc)
    gy = f(np.sqrt(f(f(1.0) + f(f(ra * rc) * f(ra * rc)))))
    err = f(f(w1 * f(f(abs(x)) * sc)) + f(w2 * gy))
    err = f(err + f(f(K_PEAR_ROUND * f(f(abs(x)) + f(abs(t)))) * sc))
    err = f(err + K_EUC_ABS)
    return float(f(r + err))


def centred(v, n):
    mean = math.fsum(v) / n
    css = math.fsum((v - mean) ** 2)
    return mean, css


@pytest.mark.parametrize("signed", [False, True])
def test_pearson_upper_bound(signed):
    rng = np.random.default_rng(21 + signed)
    nrow = 3000
    e0, e1 = constants(nrow)
    emax = e0 + e1 * nrow
    fold = 1.0 / (1.0 - emax)
    for trial in range(60):
        nnz = int(rng.integers(30, 1200))
        q, r = pair(rng, nrow, nnz, rng.random(), signed)
        if trial % 15 == 0:
            r[r != 0] = 2.5  # stored values all equal (not constant over the rows)
        x = slab_dot(q, r, rng)
        mq, cq = centred(q, nrow)
        mr, cr = centred(r, nrow)
        exact = math.fsum((q - mq) * (r - mr)) / (math.sqrt(cq) * math.sqrt(cr))
        ub = pear_upper_bound(math.sqrt(nrow) * mq, 1.0 / math.sqrt(cq), math.sqrt(nrow) * mr,
                              1.0 / math.sqrt(cr), x, nnz, e0 * fold, e1 * fold, not signed)
        assert exact <= ub, (trial, exact, ub)
        # the bound is useful: within a few hundred ppm of the value
        assert ub - exact < 2e-3, (trial, ub - exact)


def plain_constants(nrow):
    g = 1.0 / (1.0 - nrow * U23)
    return 2.0 * UB + UB * UB, (1.0 + UB) ** 2 * U23 * g


def plain_dot(q, r, rng):
    t = (bf16(q.astype(np.float32)).astype(np.float64) * bf16(r.astype(np.float32))).astype(
        np.float32)
    t = t[t != 0]
    rng.shuffle(t)
    return float(np.cumsum(t, dtype=np.float32)[-1]) if t.size else 0.0


def test_plain_bf16_bound_nonnegative():
    rng = np.random.default_rng(31)
    nrow = 4000
    e0, e1 = plain_constants(nrow)
    fold = 1.0 / (1.0 - (e0 + e1 * nrow))
    worst = 0.0
    for trial in range(60):
        nnz = int(rng.integers(50, 1500))
        q, r = pair(rng, nrow, nnz, rng.random(), False)
        x = plain_dot(q, r, rng)
        exact = math.fsum(q * r)
        bound = (e0 + e1 * nnz) * exact
        assert abs(x - exact) <= bound, trial
        worst = max(worst, abs(x - exact) / bound)
        d2 = math.fsum((q - r) ** 2)
        lb = key_lower_bound(math.fsum(q * q), math.fsum(r * r), x, nnz, e0 * fold, e1 * fold,
                             True)
        assert lb <= d2, (trial, lb, d2)
    # the representation term is attained to within a factor of a few (values
    # just past rounding ties are in every pair), so the constant is not loose
    assert worst > 0.05
