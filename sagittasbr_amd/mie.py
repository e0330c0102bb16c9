"""Monostatic RCS of a perfectly conducting sphere by the Mie series.

Textbook result (Bohren & Huffman 1983, sec. 4.4 and 8; Ruck et al. 1970):
with size parameter x = k a and spherical Bessel / Hankel functions,

    a_n = j_n(x) / h_n(x)
    b_n = [x j_n(x)]' / [x h_n(x)]'
    sigma / (pi a^2) = (1 / x^2) | sum_{n>=1} (-1)^n (2n + 1) (b_n - a_n) |^2

The series is summed to n = x + 4 x^(1/3) + 10 (Wiscombe's criterion).
Limits: sigma/(pi a^2) -> 9 x^4 for x -> 0 (Rayleigh) and -> 1 for
x -> infinity (optical). The reference ships the same closed form as a
validation script (util/monostatic_PEC_sphere.py:8-58); this is an
independent restatement.
"""
from __future__ import annotations

import numpy as np
from scipy.special import spherical_jn, spherical_yn

C0 = 299792458.0


def normalized_rcs(x):
    """sigma / (pi a^2) of a PEC sphere at size parameter(s) x = k a."""
    xs = np.atleast_1d(np.asarray(x, np.float64))
    out = np.empty_like(xs)
    for i, xi in enumerate(xs):
        nmax = int(xi + 4.0 * xi ** (1.0 / 3.0) + 10)
        n = np.arange(1, nmax + 1)
        jn = spherical_jn(n, xi)
        yn = spherical_yn(n, xi)
        djn = spherical_jn(n, xi, derivative=True)
        dyn = spherical_yn(n, xi, derivative=True)
        hn = jn + 1j * yn
        dhn = djn + 1j * dyn
        a = jn / hn
        b = (jn + xi * djn) / (hn + xi * dhn)
        s = np.sum((-1.0) ** n * (2 * n + 1) * (b - a))
        out[i] = np.abs(s) ** 2 / xi ** 2
    return out if np.ndim(x) else float(out[0])


def sphere_rcs(radius, freq):
    """Monostatic sigma [m^2] of a PEC sphere of ``radius`` [m] at ``freq`` [Hz]."""
    k = 2.0 * np.pi * np.asarray(freq, np.float64) / C0
    return normalized_rcs(k * radius) * np.pi * radius ** 2
