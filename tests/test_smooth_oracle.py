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


def spanning_line(m, L=0.1, amp=None, z_amp=None):
    """A closed loop of m points on a line along x that wraps the periodic
    box once (x_j = x0 + j L / m, stored wrapped into [0, L)), with optional
    transverse perturbations amp_y * sin(pi k j / (m - 1)) in y and z: pure
    sine modes k of the reference's odd extension of length 2m - 2."""
    j = np.arange(m)
    x0, y0, z0 = 0.013, 0.04, 0.07
    x = np.stack([x0 + j * L / m, np.full(m, y0), np.full(m, z0)], 1)
    for c, pert in ((1, amp), (2, z_amp)):
        if pert:
            for k, a in pert:
                x[:, c] += a * np.sin(np.pi * k * j / (m - 1))
    unwrapped = x.copy()
    x[:, 0] = np.mod(x[:, 0], L)
    conn = np.stack([(j + 1) % m, (j - 1) % m], 1)
    st = np.zeros((m, 2), dtype=np.int32)
    return x, conn, st, unwrapped


def smooth_one(x, conn, st, L=0.1):
    return S.smooth_tangle(x, conn, st, [0, 0, 0], [L, L, L], [1, 1, 1], 0.8)


def test_odd_extension_is_antisymmetric():
    """ext has period 2m - 2 and is odd about j = 0 and about j = m - 1:
    ext[N - j] = -ext[j], ext[0] = ext[m - 1] = 0, its first m entries are
    the detrended coordinate, and its DFT is purely imaginary."""
    m = 37
    rng = np.random.default_rng(5)
    a = np.cumsum(rng.normal(size=m)) + 3.0
    ext, coef, x0 = S.detrend_odd(a)
    N = 2 * m - 2
    assert len(ext) == N
    assert coef == (a[-1] - a[0]) / (m - 1) and x0 == a[0]
    assert ext[0] == 0.0 and ext[m - 1] == 0.0
    assert np.array_equal(ext[:m], np.concatenate([[0.0], a[1:m - 1] - coef * np.arange(1, m - 1)
                                                   - x0, [0.0]]))
    j = np.arange(1, N)
    assert np.array_equal(ext[N - j], -ext[j])
    assert np.abs(np.fft.fft(ext).real).max() < 1e-12 * np.abs(ext).sum()
    # retrend undoes detrend where nothing was filtered
    assert np.allclose(S.retrend(ext, m, coef, x0), a, rtol=0, atol=1e-14)


@pytest.mark.parametrize("m", [64, 101])
def test_straight_spanning_line_unchanged(m):
    x, conn, st, un = spanning_line(m)
    assert S.spans_box(un, [0.1] * 3)
    y = smooth_one(x, conn, st)
    assert np.abs(y - un).max() < 1e-15


def test_spanning_line_low_mode_kept_high_mode_removed():
    """N = 2m - 2 = 200: the filter zeroes halfcomplex entries 20..180, i.e.
    every sine mode k >= 20 of the odd extension; k = 3 and 7 pass."""
    m = 101
    low = [(3, 2e-3), (7, -5e-4)]
    high = [(60, 1e-3), (95, 4e-4)]
    x, conn, st, un = spanning_line(m, amp=low, z_amp=[(2, 1e-3)])
    y = smooth_one(x, conn, st)
    assert np.abs(y - un).max() < 1e-15
    x2, conn, st, un2 = spanning_line(m, amp=low + high, z_amp=[(2, 1e-3)] + high)
    y2 = smooth_one(x2, conn, st)
    assert np.abs(y2 - un).max() < 1e-15
    # the high modes were there to remove
    assert np.abs(y2 - un2).max() > 1e-4


def test_spanning_branch_idempotent():
    m = 80
    rng = np.random.default_rng(9)
    pert = [(k, a) for k, a in zip(range(1, 40), rng.normal(scale=3e-4, size=39))]
    x, conn, st, un = spanning_line(m, amp=pert, z_amp=pert[::-1])
    y = smooth_one(x, conn, st)
    y[:, 0] = np.mod(y[:, 0], 0.1)
    z = smooth_one(y, conn, st)
    z[:, 0] = np.mod(z[:, 0], 0.1)
    assert np.abs(z - y).max() < 1e-14


def test_circle_loop_unchanged():
    n, R = 120, 0.02
    ph = 2 * np.pi * np.arange(n) / n
    x = np.stack([0.05 + R * np.cos(ph), 0.05 + R * np.sin(ph), np.full(n, 0.05)], 1)
    conn = np.stack([(np.arange(n) + 1) % n, (np.arange(n) - 1) % n], 1)
    st = np.zeros((n, 2), dtype=np.int32)
    y = S.smooth_tangle(x, conn, st, [0, 0, 0], [0.1, 0.1, 0.1], [1, 1, 1], 0.8)
    assert np.abs(y - x).max() < 1e-15
