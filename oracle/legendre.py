"""Normalised associated Legendre functions lambda_lm(cos theta), FP64 on the CPU.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

lambda_lm = N_lm * P_lm with N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)
(``sht_raw.py:26-28`` / ``sht_fallback.py:4-8``) and P_lm the unnormalised
function with the Condon-Shortley phase (``legendre.cu:13-21``).

Reference recursion (``legendre.cu:13-34``) overflows at m >= 151 and the
host norm raises for l+m >= 171; this module evaluates the same functions with
the textbook normalised recursion

    lambda_mm   = (-1)^m sqrt((2m+1)/(4 pi) prod_{k=1..m} (2k-1)/(2k)) sin^m theta
    lambda_l,m  = a_lm (x lambda_{l-1,m} - lambda_{l-2,m} / a_{l-1,m}),
    a_lm        = sqrt((4 l^2 - 1) / (l^2 - m^2))

carried as mantissa * 2**E so that sin^m theta never underflows.
"""
import math
from fractions import Fraction

import numpy as np

_BIG = 2.0 ** 600
_BIG_E = 600


def lambda_mm_scaled(mmax, sin_theta):
    """Yield (m, mantissa, exponent) with lambda_mm = mantissa * 2**exponent.

    Iterates lambda_mm = -sqrt((2m+1)/(2m)) sin(theta) lambda_{m-1,m-1},
    lambda_00 = 1/sqrt(4 pi), renormalising with frexp every step.
    """
    st = np.asarray(sin_theta, dtype=np.float64)
    mant = np.full(st.shape, 1.0 / math.sqrt(4.0 * math.pi))
    expo = np.zeros(st.shape, dtype=np.int64)
    yield 0, mant.copy(), expo.copy()
    for m in range(1, mmax + 1):
        mant = mant * (-math.sqrt((2.0 * m + 1.0) / (2.0 * m))) * st
        fm, fe = np.frexp(mant)
        mant = fm
        expo = expo + fe
        yield m, mant.copy(), expo.copy()


def lambda_block(m, lmax, x, mant_mm, exp_mm):
    """lambda_lm(x) for l = m..lmax as an array [lmax-m+1, len(x)] (FP64).

    Values below the FP64 range flush to 0 exactly as their true magnitude
    (< 2^-1074) dictates.
    """
    x = np.asarray(x, dtype=np.float64)
    n = lmax - m + 1
    out = np.zeros((n, x.size), dtype=np.float64)
    p1 = mant_mm.astype(np.float64).copy()      # lambda_{l-1} mantissa
    p2 = np.zeros_like(p1)                      # lambda_{l-2} mantissa
    e = exp_mm.astype(np.int64).copy()
    out[0] = np.ldexp(p1, e)
    a_prev = None
    for l in range(m + 1, lmax + 1):
        a_l = math.sqrt((4.0 * l * l - 1.0) / float((l - m) * (l + m)))
        if a_prev is None:
            new = a_l * x * p1
        else:
            new = a_l * (x * p1 - p2 / a_prev)
        big = np.abs(new) > _BIG
        if big.any():
            new = np.where(big, new / _BIG, new)
            p1 = np.where(big, p1 / _BIG, p1)
            e = np.where(big, e + _BIG_E, e)
        p2, p1 = p1, new
        out[l - m] = np.ldexp(p1, e)
        a_prev = a_l
    return out


def lambda_all(lmax, mmax, cos_theta, sin_theta):
    """Yield (m, block) for m = 0..mmax with block = lambda_block(m, ...)."""
    for m, mant, ex in lambda_mm_scaled(mmax, sin_theta):
        yield m, lambda_block(m, lmax, cos_theta, mant, ex)


# ---------------------------------------------------------------------------
# Literal reference restatements (valid where the reference is finite)
# ---------------------------------------------------------------------------

def plm_literal(x, lmax, m):
    """Unnormalised P_l^m(x), l=0..lmax, as ``legendre.cu:2-35`` computes it.

    Output [len(x), lmax+1] with columns < m zero.  FP64; overflows at
    m >= 151 exactly as the reference does.
    """
    x = np.asarray(x, dtype=np.float64)
    P = np.zeros((x.size, lmax + 1))
    somx2 = np.sqrt(np.maximum(0.0, (1.0 - x) * (1.0 + x)))   # legendre.cu:10
    pmm = np.ones_like(x)
    fact = 1.0
    for _ in range(1, m + 1):                                 # legendre.cu:14-19
        pmm = pmm * (-fact * somx2)
        fact += 2.0
    P[:, m] = pmm
    if m == lmax:
        return P
    pmmp1 = x * (2.0 * m + 1.0) * pmm                         # legendre.cu:23
    P[:, m + 1] = pmmp1
    p2, p1 = pmm, pmmp1
    for l in range(m + 2, lmax + 1):                          # legendre.cu:27-34
        pl = ((2.0 * l - 1.0) * x * p1 - (l + m - 1.0) * p2) / float(l - m)
        P[:, l] = pl
        p2, p1 = p1, pl
    return P


def alm_norm_exact(l, m):
    """``_alm_norm`` (sht_raw.py:26-28) with the factorial ratio taken exactly.

    The reference's ``float(factorial(l+m))`` raises for l+m >= 171; the ratio
    (l-m)!/(l+m)! is formed as an exact rational first.
    """
    ratio = Fraction(math.factorial(l - m), math.factorial(l + m))
    return math.sqrt((2 * l + 1) / (4.0 * math.pi) * float(ratio))
