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


def open_line(m, z_mode=None, amp=0.0, k=0):
    """m points from the DOWN wall (z = 0) to the UP wall (z = 0.1) at
    x = y = 0.05, slots in order, both ends pinned; optionally x displaced
    by amp sin(pi k j / (m - 1)) (a mode of the odd extension)."""
    j = np.arange(m)
    x = np.stack([np.full(m, 0.05), np.full(m, 0.05), 0.1 * j / (m - 1)], 1)
    if amp:
        x[:, 0] += amp * np.sin(np.pi * k * j / (m - 1))
    conn = np.stack([np.where(j < m - 1, j + 1, -1), np.where(j > 0, j - 1, -1)], 1)
    st = np.zeros((m, 2), dtype=np.int64)
    st[0] = (1, 4)
    st[-1] = (1, 5)
    return x, conn, st


def smooth_lines(x, conn, st, fcut=0.25):
    return S.smooth_tangle(x, conn, st, [0, 0, 0], [0.1, 0.1, 0.1], [1, 1, 0], fcut)


def test_open_line_straight_unchanged():
    x, conn, st = open_line(40)
    y = smooth_lines(x, conn, st)
    assert np.abs(y - x).max() < 1e-16


def test_open_line_low_mode_kept_high_mode_removed():
    """The open branch (src/vortex_smooth.c:316-444) filters the odd
    extension of length 2m - 2, whose modes are sin(pi k j / (m - 1)): a low
    k passes unchanged, one in the cut band is removed; the end points stay
    exactly where they were."""
    m = 33
    x0, conn, st = open_line(m)
    lo, c1, s1 = open_line(m, amp=1e-3, k=2)
    hi, _, _ = open_line(m, amp=1e-3, k=m - 3)
    ylo = smooth_lines(lo, c1, s1)
    yhi = smooth_lines(hi, c1, s1)
    assert np.abs(ylo - lo).max() < 1e-15
    assert np.abs(yhi - x0).max() < 1e-15
    for y, src in ((ylo, lo), (yhi, hi)):
        assert np.array_equal(y[[0, m - 1]], src[[0, m - 1]])


def test_open_line_is_the_spanning_branch_with_fixed_ends():
    """An open line is filtered as a box-spanning loop of the same points is
    (detrend, odd extension, low pass, retrend), except that its end points
    are not written; the line is found from a mid-line slot by its reverse
    links (is_vortex_open, :24-39)."""
    rng = np.random.default_rng(4)
    m = 21
    x, conn, st = open_line(m)
    x[1:-1] += rng.normal(scale=2e-4, size=(m - 2, 3))
    perm = np.roll(np.arange(m), 7)  # slot s holds line point perm[s]
    inv = np.argsort(perm)
    xs = x[perm]
    cs = np.stack([np.where(conn[perm, 0] >= 0, inv[np.maximum(conn[perm, 0], 0)], -1),
                   np.where(conn[perm, 1] >= 0, inv[np.maximum(conn[perm, 1], 0)], -1)], 1)
    y = smooth_lines(xs, cs, st[perm])[inv]
    want = np.zeros_like(x)
    for c in range(3):
        ext, coef, x0 = S.detrend_odd(x[:, c])
        want[:, c] = S.retrend(S.low_pass(ext, 0.25), m, coef, x0)
    assert np.abs(y[1:-1] - want[1:-1]).max() < 1e-16
    assert np.array_equal(y[[0, m - 1]], x[[0, m - 1]])
