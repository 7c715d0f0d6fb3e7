"""Numpy restatement of the reference's frequency-filter smoothing.

TEST INFRASTRUCTURE ONLY (the checker for ovg_smooth).

Follows smooth -> high_freq_cutoff -> low_pass_filter
(src/vortex_smooth.c:134-175,183-453,459-462) for closed loops and for
open lines ending on walls.  The
reference calls FFTW (fftw_plan_r2r_1d, :142,163), a third-party library
that is not vendored and not installed here; the thesis names FFTW 3.3.9
(section 1.3.1 p.11).  Its halfcomplex R2HC layout is restated with
numpy.fft.rfft: r_k = Re X_k, i_k = Im X_k with X_k = sum_j x_j e^{-2 pi i jk/N};
the band of entries the reference zeroes is removed and the halfcomplex
inverse (HC2R) is taken, then divided by N as the reference does.

The forward plan receives FFTW_FORWARD (-1) as its r2r kind (:142) and the
backward plan FFTW_BACKWARD (+1, :163).  In FFTW 3's fftw_r2r_kind
enumeration +1 is FFTW_HC2R (R2HC = 0, HC2R = 1, DHT = 2, REDFT00 = 3, ...),
so the backward plan is the halfcomplex inverse; -1 is no r2r kind at all.
This restatement takes the forward plan as R2HC, the transform whose
inverse HC2R is and the layout the reference's own comment at :148
describes.  That reading cannot be checked without FFTW: FFTW-level parity
of the smoothing is unpinned.  The restatement itself (the filter, the
detrend and odd extension of box-spanning loops) is pinned by closed forms
in tests/test_smooth_oracle.py.
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


def spans_box(xs, L):
    """The reference's test for a loop that wraps the box: its unwrapped
    start and end points more than half a box width apart along x's box
    side (src/vortex_smooth.c:246-247)."""
    return np.sqrt(np.sum((xs[0] - xs[-1]) ** 2)) > L[0] / 2


def detrend_odd(a):
    """One coordinate of a box-spanning loop (src/vortex_smooth.c:248-285):
    the straight line from the first to the last point is subtracted (both
    ends become 0) and the result is continued to length 2m - 2 by its odd
    reflection about the last point, ext[m + j] = -a[m - 2 - j].  Returns
    (ext, coef, x0) with the line x0 + coef * j."""
    a = np.array(a, dtype=np.float64)
    m = len(a)
    coef = (a[m - 1] - a[0]) / (m - 1)
    x0 = a[0]
    a[1:m - 1] = a[1:m - 1] - coef * np.arange(1, m - 1) - x0
    a[0] = a[m - 1] = 0.0
    return np.concatenate([a, -a[m - 2:0:-1]]), coef, x0


def retrend(f, m, coef, x0):
    """The first m filtered values with the line added back
    (src/vortex_smooth.c:287-292)."""
    return f[:m] + coef * np.arange(m) + x0


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


def line_start(conn, i):
    """is_vortex_open (src/vortex_smooth.c:24-39): walking the reverse links
    from i, the point whose reverse link is -1 for an open line (its
    start), None for a closed loop."""
    k = i
    while conn[k, 1] != i and conn[k, 1] != -1:
        k = conn[k, 1]
    return k if conn[k, 1] == -1 else None


def smooth_tangle(vnodes, conn, status, box_lo, box_hi, periodic, freq_to_cutoff):
    """high_freq_cutoff (src/vortex_smooth.c:183-453); returns new positions
    (unwrapped along each loop, as the reference leaves them before
    enforce_periodic_boundaries).  Loops are visited from their smallest
    slot; a closed loop is listed forward from there, an open line (ends
    pinned to walls) forward from its start (:316-352).  Open lines are
    always detrended and odd-extended (:354-381, as box-spanning loops are,
    :245-292), filtered at length 2m - 2 and retrended (:421-433), and keep
    their two end points (:435-444)."""
    vnodes = np.array(vnodes, dtype=np.float64)
    n = len(vnodes)
    L = np.asarray(box_hi, dtype=np.float64) - np.asarray(box_lo, dtype=np.float64)
    seen = np.zeros(n, dtype=bool)
    out = vnodes.copy()
    for i in range(n):
        if seen[i] or status[i, 0] == -1:
            continue
        start = line_start(conn, i)
        is_open = start is not None
        slots = []
        k = start if is_open else i
        while True:
            seen[k] = True
            slots.append(k)
            k = conn[k, 0]
            if k == slots[0] or k == -1:
                break
        m = len(slots)
        xs = np.zeros((m, 3))
        xs[0] = vnodes[slots[0]]
        for j in range(1, m):
            xs[j] = pwrap_next(xs[j - 1], vnodes[slots[j]], L, periodic)
        if is_open and m < 3:
            continue  # no interior point to move
        span = is_open or spans_box(xs, L)
        ys = np.zeros_like(xs)
        for c in range(3):
            if span:
                ext, coef, x0 = detrend_odd(xs[:, c])
                ys[:, c] = retrend(low_pass(ext, freq_to_cutoff), m, coef, x0)
            else:
                ys[:, c] = low_pass(xs[:, c], freq_to_cutoff)
        if is_open:
            out[slots[1:m - 1]] = ys[1:m - 1]
        else:
            out[slots] = ys
    return out
