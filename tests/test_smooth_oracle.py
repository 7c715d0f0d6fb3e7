"""Pins of the numpy smoothing restatement (oracle/smooth.py).

FFTW (the reference's smoothing arithmetic, src/vortex_smooth.c:138-174) is
absent, so FFTW-level parity is unpinned; the restatement is pinned by an
explicit O(N^2) halfcomplex DFT that zeroes the same index band as the C
loops (src/vortex_smooth.c:146-160), and by closed forms.
"""
import numpy as np
import pytest

from oracle import smooth as S


def explicit_filter(x, fcut):
    """R2HC by definition, zero out[idx] for the C loop's idx range, HC2R by
    definition, divide by N."""
    N = len(x)
    j = np.arange(N)
    out = np.zeros(N)
    for k in range(N // 2 + 1):
        out[k] = np.sum(x * np.cos(2 * np.pi * j * k / N))
    for k in range(1, (N - 1) // 2 + 1):
        out[N - k] = -np.sum(x * np.sin(2 * np.pi * j * k / N))
    s = int(fcut * N / 2)
    if N > 2 * s + 1:
        rng = range(N // 2 - s, N // 2 + 1 + s) if N % 2 == 0 else range(N // 2 - s, N // 2 + s)
        for i in rng:
            out[i] = 0
    y = np.full(N, out[0])
    for k in range(1, (N - 1) // 2 + 1):
        y += 2 * (out[k] * np.cos(2 * np.pi * j * k / N) - out[N - k] * np.sin(2 * np.pi * j * k / N))
    if N % 2 == 0:
        y += out[N // 2] * np.cos(np.pi * j)
    return y / N


@pytest.mark.parametrize("N", [5, 6, 7, 8, 31, 32, 33, 100, 101])
@pytest.mark.parametrize("fcut", [0.0, 0.3, 0.8])
def test_low_pass_matches_explicit_halfcomplex(N, fcut):
    x = np.random.default_rng(N).normal(size=N)
    assert np.allclose(S.low_pass(x, fcut), explicit_filter(x, fcut), atol=1e-12)


def test_low_modes_pass_high_modes_removed():
    N = 200
    j = np.arange(N)
    low = np.cos(2 * np.pi * 3 * j / N) + 0.5 * np.sin(2 * np.pi * 5 * j / N) + 2.0
    high = np.cos(2 * np.pi * 90 * j / N)
    assert np.allclose(S.low_pass(low + high, 0.8), low, atol=1e-12)


def test_idempotent():
    x = np.random.default_rng(3).normal(size=64)
    y = S.low_pass(x, 0.8)
    assert np.allclose(S.low_pass(y, 0.8), y, atol=1e-13)


def test_circle_loop_unchanged():
    n, R = 120, 0.02
    ph = 2 * np.pi * np.arange(n) / n
    x = np.stack([0.05 + R * np.cos(ph), 0.05 + R * np.sin(ph), np.full(n, 0.05)], 1)
    conn = np.stack([(np.arange(n) + 1) % n, (np.arange(n) - 1) % n], 1)
    st = np.zeros((n, 2), dtype=np.int32)
    y = S.smooth_tangle(x, conn, st, [0, 0, 0], [0.1, 0.1, 0.1], [1, 1, 1], 0.8)
    assert np.abs(y - x).max() < 1e-15
