"""Numpy restatement of the reference's frequency-filter smoothing.

TEST INFRASTRUCTURE ONLY (the checker for ovg_smooth).

Follows smooth -> high_freq_cutoff -> low_pass_filter
(src/vortex_smooth.c:134-175,183-453,459-462) for closed loops.  The
reference calls FFTW (fftw_plan_r2r_1d, :142,163), a third-party library
that is not vendored and not installed here; the thesis names FFTW 3.3.9
(section 1.3.1 p.11).  Its halfcomplex R2HC layout is restated with
numpy.fft.rfft: r_k = Re X_k, i_k = Im X_k with X_k = sum_j x_j e^{-2 pi i jk/N};
the band of entries the reference zeroes is removed and the halfcomplex
inverse (HC2R) is taken, then divided by N as the reference does.

The forward plan receives FFTW_FORWARD (-1) as its r2r kind (:142); this
restatement takes it as R2HC, the layout the reference's own comment at
:148 describes.  That reading cannot be checked without FFTW: FFTW-level
parity of the smoothing is unpinned.  The restatement itself is pinned by
closed forms in tests/test_smooth_oracle.py.
"""
import numpy as np


def kept_masks(N, freq_to_cutoff):
    """Boolean masks over k = 0..N//2 of the real and imaginary halfcomplex
    entries that survive (src/vortex_smooth.c:146-160)."""
    half = N // 2
    keep_r = np.ones(half + 1, dtype=bool)
    keep_i = np.ones(half + 1, dtype=bool)
    keep_i[0] = False
    if N % 2 == 0:
        keep_i[half] = False  # no imaginary Nyquist entry
    s = int(freq_to_cutoff * N / 2)
    if N > 2 * s + 1:
        if N % 2 == 0:
            zero = range(half - s, half + 1 + s)
        else:
            zero = range(half - s, half + s)
        for idx in zero:
            if idx <= half:
                keep_r[idx] = False
            else:
                keep_i[N - idx] = False
    return keep_r, keep_i


def low_pass(x, freq_to_cutoff):
    """low_pass_filter of one coordinate array (src/vortex_smooth.c:134-175)."""
    x = np.asarray(x, dtype=np.float64)
    N = len(x)
    X = np.fft.rfft(x)
    kr, ki = kept_masks(N, freq_to_cutoff)
    Y = np.where(kr, X.real, 0.0) + 1j * np.where(ki, X.imag, 0.0)
    # HC2R of the kept entries; irfft normalises by 1/N already, which is the
    # reference's division by N after the unnormalised backward transform
    return np.fft.irfft(Y, n=N)


def pwrap_next(prev, cur, L, periodic):
    """seg_pwrap of cur against prev (src/boundaries.c:54-83)."""
    out = np.array(cur, dtype=np.float64)
    for a in range(3):
        if not periodic[a]:
            continue
        while (out[a] - prev[a]) > L[a] / 2:
            out[a] -= L[a]
        while (out[a] - prev[a]) < -L[a] / 2:
            out[a] += L[a]
    return out


def smooth_tangle(vnodes, conn, status, box_lo, box_hi, periodic, freq_to_cutoff):
    """high_freq_cutoff for closed loops; returns new positions (unwrapped
    along each loop, as the reference leaves them before
    enforce_periodic_boundaries)."""
    vnodes = np.array(vnodes, dtype=np.float64)
    n = len(vnodes)
    L = np.asarray(box_hi, dtype=np.float64) - np.asarray(box_lo, dtype=np.float64)
    seen = np.zeros(n, dtype=bool)
    out = vnodes.copy()
    for i in range(n):
        if seen[i] or status[i, 0] == -1:
            continue
        slots = []
        k = i
        while True:
            seen[k] = True
            slots.append(k)
            k = conn[k, 0]
            if k == i:
                break
        m = len(slots)
        xs = np.zeros((m, 3))
        xs[0] = vnodes[slots[0]]
        for j in range(1, m):
            xs[j] = pwrap_next(xs[j - 1], vnodes[slots[j]], L, periodic)
        span = np.sqrt(np.sum((xs[0] - xs[m - 1]) ** 2)) > L[0] / 2
        ys = np.zeros_like(xs)
        for c in range(3):
            a = xs[:, c].copy()
            if span:
                coef = (a[m - 1] - a[0]) / (m - 1)
                x0 = a[0]
                a[1:m - 1] = a[1:m - 1] - coef * np.arange(1, m - 1) - x0
                a[0] = a[m - 1] = 0.0
                ext = np.concatenate([a, -a[m - 2:0:-1]])
                f = low_pass(ext, freq_to_cutoff)[:m]
                ys[:, c] = f + coef * np.arange(m) + x0
            else:
                ys[:, c] = low_pass(a, freq_to_cutoff)
        out[slots] = ys
    return out
