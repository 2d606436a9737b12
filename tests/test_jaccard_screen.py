"""The one-FMA Jaccard screen (jac_screen_coef in
sparsedist_amd/csrc/k_slab.hip), checked against an emulation of its FP32
arithmetic on the CPU.  The filtered pass screens against each query's
threshold with it, and the list-mode pass marks the values it visits with it
against the list's current L-th key (an exact rational whose FP64 quotient is
the T below).

A candidate with intersection c of a reference column with ni stored entries
and a query with nj enters the buffers when its exact key, the FP64 quotient
c / (ni + nj - c), is at least the query's FP64 threshold T (1 for two empty
columns).  The kernel first screens it in FP32: the screen threshold is
t = fl(T) - |fl(T)| 2^-20 - 1e-30, the coefficient a = (max(t, 0) /
(1 + max(t, 0))) (1 - 2^-18) (each step rounded to FP32), b = fl(a nj), and the
value passes when c >= fma(a, ni, b).  The screen may keep too much but must
never drop a candidate the exact test keeps.  The tightest case is a
threshold equal to the candidate's own key, which the checks below enumerate;
the exact side uses Python fractions, and the FMA's single rounding is
emulated by rounding the exact value a ni + b to FP32 once.
"""
from fractions import Fraction

import numpy as np

F32 = np.float32


def f32(x):
    """Round to the nearest FP32 (ties to even).  The values rounded here,
    a ni + b with a, b FP32 and ni < 2^24, carry at most 48 significant bits,
    so the conversion to FP64 is exact and the one rounding is to FP32."""
    return F32(float(x))


def coef(t32):
    tt = max(t32, F32(0.0))
    q = F32(tt / F32(F32(1.0) + tt))
    return F32(q * F32(0.999996185302734375))


def screen_value(T, ni, nj):
    ft = F32(T)
    t = F32(F32(ft - F32(abs(ft) * F32(9.5367431640625e-07))) - F32(1e-30))
    a = coef(t)
    if np.isnan(a):  # no threshold (an infinite one): the FMA is NaN
        return F32(np.nan)
    b = F32(a * F32(nj))
    exact = Fraction(float(a)) * ni + Fraction(float(b))  # fma: one rounding
    return f32(exact)


def exact_key(c, ni, nj):
    d = ni + nj - c
    return 1.0 if d == 0 else c / d  # FP64 quotient, correctly rounded


def test_screen_keeps_every_candidate_at_its_own_key():
    rng = np.random.default_rng(5)
    cases = 0
    for _ in range(60000):
        ni = int(rng.integers(0, 30000))
        nj = int(rng.integers(0, 30000))
        c = int(rng.integers(0, min(ni, nj) + 1))
        T = exact_key(c, ni, nj)
        assert F32(c) >= screen_value(T, ni, nj), (c, ni, nj, T)
        cases += 1
    assert cases == 60000


def test_screen_keeps_candidates_near_the_boundary():
    """Thresholds from a neighbouring rational just below and at the key."""
    rng = np.random.default_rng(6)
    for _ in range(20000):
        ni = int(rng.integers(1, 20000))
        nj = int(rng.integers(1, 20000))
        c = int(rng.integers(1, min(ni, nj) + 1))
        key = exact_key(c, ni, nj)
        for T in (key, np.nextafter(key, 0.0), np.nextafter(key, 2.0)):
            if key >= T:  # the exact test keeps it
                assert F32(c) >= screen_value(T, ni, nj), (c, ni, nj, T)


def test_edge_cases():
    # two empty columns: J = 1 passes any threshold up to 1
    assert F32(0.0) >= screen_value(1.0, 0, 0)
    # threshold 0 keeps everything, including c = 0
    assert F32(0.0) >= screen_value(0.0, 100, 200)
    # no query (infinite threshold): nothing passes
    assert not (F32(5.0) >= screen_value(np.inf, 10, 10))
    # large counts just below 2^24 stay exact in FP32
    ni, nj = (1 << 23) + 11, (1 << 23) - 7
    c = nj
    assert F32(c) >= screen_value(exact_key(c, ni, nj), ni, nj)


def test_screen_check_has_teeth():
    """A screen without the slack (coefficient T / (1 + T) from the exact
    threshold, no lowering) drops some candidates at their own key in the
    same emulation, so the lowering is what the kernel's guarantee rests on."""
    rng = np.random.default_rng(7)
    dropped = 0
    for _ in range(60000):
        ni = int(rng.integers(1, 30000))
        nj = int(rng.integers(1, 30000))
        c = int(rng.integers(1, min(ni, nj) + 1))
        ft = F32(exact_key(c, ni, nj))
        a0 = F32(ft / F32(F32(1.0) + ft))
        b0 = F32(a0 * F32(nj))
        dropped += int(F32(c) < f32(Fraction(float(a0)) * ni + Fraction(float(b0))))
    assert dropped > 0
