// Frequency-filter smoothing (smooth -> high_freq_cutoff -> low_pass_filter,
// src/vortex_smooth.c:134-175,183-453,459-462): one workgroup per loop up
// to kSmallLoop points, one batched Bluestein transform over all longer
// loops.
//
// Per loop (in slot order of its first slot, src/vortex_smooth.c:192-210)
// the coordinates are unwrapped along the loop, each point within half a box
// of the previous unwrapped one (:229-242).  A loop whose ends are more than half
// a box apart spans the box; it is detrended and odd-extended to length
// 2n-2 (:245-292).  Each coordinate array of length N is then low-passed:
// the reference takes an FFTW r2r transform (halfcomplex r_k, i_k), zeroes
// the halfcomplex entries N/2-s .. N/2+s (even N) or N/2-s .. N/2+s-1
// (odd N) with s = (int)(freq_to_cutoff*N/2) when N > 2s+1 (:146-160),
// transforms back and divides by N (:163-169).  The filter is linear, so it
// is evaluated here directly: the kept coefficients
//   r_k = sum_j x_j cos(2 pi jk/N),  i_k = -sum_j x_j sin(2 pi jk/N)
// followed by the halfcomplex inverse
//   y_j = [r_0 + 2 sum_k (r_k cos - i_k sin) + r_{N/2} (-1)^j] / N
// restricted to the entries that survive.  Twiddles use exact integer
// argument reduction (jk mod N) and sincospi.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "engine_internal.h"

namespace ovg {

constexpr int kSmBlock = 256;
constexpr int kSmallLoop = 256;      // DFT of the kept band, 256 threads per loop

// ---- shared pieces -----------------------------------------------------------

// Inclusive block scan of one int per thread (Hillis-Steele in LDS).
__device__ int block_scan_incl(int v, int *sh)
{
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        int a = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    int r = sh[t];
    __syncthreads();
    return r;
}

// Whole periods that seg_pwrap adds to the link difference d = x_j - x_{j-1}
// along axis a (src/vortex_smooth.c:229-242).
__device__ __forceinline__ int link_wraps(const TangleParams &P, int a, double d)
{
    int w = 0, g = 0;
    while (d > P.halfL[a] && g++ < 64) { d -= P.L[a]; --w; }
    while (d < -P.halfL[a] && g++ < 128) { d += P.L[a]; ++w; }
    return w;
}

// Unwrap a loop's coordinates along its links: x'_0 = x_0 and x'_j is x_j
// moved by whole periods to within half a box of x'_{j-1}
// (src/vortex_smooth.c:229-242).  The period count of x_j is the running
// sum of the per-link wraps that seg_pwrap applies to x_j - x_{j-1}, so the
// whole block computes it with a prefix sum instead of a serial walk.
__device__ void block_unwrap(const TangleParams &P, const int *slots, int n, const double *X,
                             const double *Y, const double *Z, double *const xs[3], int *sh,
                             int *sh_carry)
{
    const double *src[3] = {X, Y, Z};
    for (int a = 0; a < 3; ++a) {
        if (threadIdx.x == 0) *sh_carry = 0;
        __syncthreads();
        for (int base = 0; base < n; base += blockDim.x) {
            int j = base + threadIdx.x;
            int w = 0;
            double cur = 0;
            if (j < n) {
                cur = src[a][slots[j]];
                if (j > 0 && P.periodic[a]) w = link_wraps(P, a, cur - src[a][slots[j - 1]]);
            }
            int incl = block_scan_incl(w, sh);
            int carry = *sh_carry;
            if (j < n) xs[a][j] = cur + (double)(carry + incl) * P.L[a];
            __syncthreads();
            if (threadIdx.x == blockDim.x - 1) *sh_carry = carry + incl;
            __syncthreads();
        }
    }
}

// Loops whose ends are more than half a box apart span the box
// (src/vortex_smooth.c:245-246).
__device__ bool loop_spans(const TangleParams &P, double *const xs[3], int n)
{
    double ex = xs[0][0] - xs[0][n - 1], ey = xs[1][0] - xs[1][n - 1],
           ez = xs[2][0] - xs[2][n - 1];
    return sqrt(dot3(ex, ey, ez, ex, ey, ez)) > (P.hi[0] - P.lo[0]) / 2;
}

// Subtract the end-to-end line and odd-extend to length 2n-2
// (src/vortex_smooth.c:248-279).
__device__ void block_detrend(double *const xs[3], int n, double *coef, double *x0)
{
    if (threadIdx.x < 3) {
        double *a = xs[threadIdx.x];
        coef[threadIdx.x] = (a[n - 1] - a[0]) / (n - 1);
        x0[threadIdx.x] = a[0];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        for (int c = 0; c < 3; ++c) {
            if (j == 0 || j == n - 1) continue;
            xs[c][j] = xs[c][j] - coef[c] * j - x0[c];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        xs[threadIdx.x][0] = 0;
        xs[threadIdx.x][n - 1] = 0;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n - 2; j += blockDim.x)
        for (int c = 0; c < 3; ++c) xs[c][n + j] = -xs[c][n - 2 - j];
    __syncthreads();
}

// (src/vortex_smooth.c:286-291) then write back along the loop (:301-309)
__device__ void block_finish(double *const xs[3], int n, bool span, bool open,
                             const double *coef, const double *x0, const int *slots, double *X,
                             double *Y, double *Z)
{
    if (span) {
        for (int j = threadIdx.x; j < n; j += blockDim.x)
            for (int c = 0; c < 3; ++c) xs[c][j] = xs[c][j] + coef[c] * j + x0[c];
        __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        if (open && (j == 0 || j == n - 1)) continue;  // the ends stay (:435-444)
        int m = slots[j];
        X[m] = xs[0][j];
        Y[m] = xs[1][j];
        Z[m] = xs[2][j];
    }
}

// Surviving halfcomplex entries for length N (src/vortex_smooth.c:146-160):
// r_k kept for 0 <= k < kr, i_k for 1 <= k < ki; false when N <= 2s+1
// (nothing is zeroed).
__device__ bool kept_band(int N, double fcut, int &kr, int &ki)
{
    const int s = (int)(fcut * N / 2);
    if (!(N > 2 * s + 1 && N >= 2)) return false;
    const int half = N / 2;
    kr = half - s;
    ki = (N % 2 == 0) ? half - s : half + 2 - s;
    if (ki > half + 1) ki = half + 1;
    if (N % 2 == 0 && ki > half) ki = half;
    return true;
}

// ---- small loops: DFT of the kept band -------------------------------------
__global__ __launch_bounds__(kSmBlock) void k_smooth(
    int n_loops, const int *__restrict__ loop_off, const int *__restrict__ loop_slots,
    const int *__restrict__ loop_open, double *__restrict__ X, double *__restrict__ Y,
    double *__restrict__ Z, const TangleParams *__restrict__ Pp, double fcut,
    double *__restrict__ work)
{
    int lp = blockIdx.x;
    if (lp >= n_loops) return;
    const TangleParams &P = *Pp;
    const int off = loop_off[lp];
    const int n = loop_off[lp + 1] - off;
    if (n > kSmallLoop) return;  // FFT paths
    // an open line (ends pinned to walls, src/vortex_smooth.c:316-444) is
    // detrended and odd-extended like a box-spanning loop and keeps its end
    // points; one of fewer than 3 points has nothing to move
    const bool open = loop_open[lp] != 0;
    if (open && n < 3) return;
    const int *slots = loop_slots + off;
    // per loop 15n + 6 doubles from 15*off + 6*lp: coordinates 3 x 2n,
    // coefficients 3 x (2n + 2), output 3 x n
    double *w = work + (size_t)15 * off + (size_t)6 * lp;
    double *xs[3] = {w, w + 2 * (size_t)n, w + 4 * (size_t)n};
    double *cf = w + 6 * (size_t)n;
    double *yo = cf + 3 * (2 * (size_t)n + 2);
    const size_t cstride = 2 * (size_t)n + 2;
    __shared__ int sh[kSmBlock];
    __shared__ int s_carry, s_span;
    __shared__ double s_coef[3], s_x0[3];
    block_unwrap(P, slots, n, X, Y, Z, xs, sh, &s_carry);
    if (threadIdx.x == 0) s_span = open || loop_spans(P, xs, n);
    __syncthreads();
    const bool span = s_span;
    if (span) block_detrend(xs, n, s_coef, s_x0);
    const int N = span ? 2 * n - 2 : n;
    int kr, ki;
    if (kept_band(N, fcut, kr, ki)) {
        const int nimag = ki > 1 ? ki - 1 : 0;
        const int ncoef = kr + nimag;
        for (int idx = threadIdx.x; idx < 3 * ncoef; idx += kSmBlock) {
            int c = idx / ncoef, q = idx % ncoef;
            const double *a = xs[c];
            bool is_r = q < kr;
            long kk = is_r ? q : q - kr + 1;
            double acc = 0;
            for (int j = 0; j < N; ++j) {
                long m = (kk * j) % N;
                double sn, cs;
                sincospi(2.0 * (double)m / N, &sn, &cs);
                acc += a[j] * (is_r ? cs : -sn);
            }
            cf[c * cstride + q] = acc;
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < 3 * n; idx += kSmBlock) {
            int c = idx / n, j = idx % n;
            const double *co = cf + c * cstride;
            double y = co[0];
            for (int q = 1; q < kr; ++q) {
                long m = ((long)q * j) % N;
                double sn, cs;
                sincospi(2.0 * (double)m / N, &sn, &cs);
                y += 2 * co[q] * cs;
            }
            for (int q = kr; q < ncoef; ++q) {
                long kk = q - kr + 1;
                long m = (kk * j) % N;
                double sn, cs;
                sincospi(2.0 * (double)m / N, &sn, &cs);
                y -= 2 * co[q] * sn;
            }
            yo[c * (size_t)n + j] = y / N;
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < 3 * n; idx += kSmBlock) {
            int c = idx / n, j = idx % n;
            xs[c][j] = yo[c * (size_t)n + j];
        }
        __syncthreads();
    }
    block_finish(xs, n, span, open, s_coef, s_x0, slots, X, Y, Z);
}

// ---- loops longer than kSmallLoop: one batched Bluestein transform ---------
// U_k = sum_j u_j e^{-2 pi i jk/N} = c_k (a * b)_k with the chirp
// c_j = e^{-pi i j^2/N}, a_j = u_j c_j, b_m = conj(c_m) (Bluestein 1970).
// The cyclic convolution of length M = 2^ceil(log2(2N-1)) is the inverse
// transform of DIF(a) . DIF(b): the decimation-in-frequency forward transform
// leaves its output in bit-reversed order, the product is taken in that
// order and the decimation-in-time inverse returns to natural order, so no
// bit-reversal pass is needed.  The filter zeroes the same halfcomplex
// entries as the small-loop path and y = conj(DFT(conj(U)))/N.
//
// All loops longer than kSmallLoop are transformed together: the three
// coordinates of every loop are signals of their own and every launch covers
// every signal, a block finding its loop through a host-built chunk map.
// Stages of span >= kTile run as grid-wide passes of up to four radix-2
// stages each (16-element groups in registers); the stages of span < kTile run as one pass over kTile-element tiles
// in LDS, which for the coordinate signals also takes the product with
// DIF(b) and the first inverse stages.

constexpr int kTileLog = 12;
constexpr int kTile = 1 << kTileLog;
constexpr int kChunk = 1024;       // points per block in the unwrap and the write-back
constexpr int kPassBlock = 256;    // threads per block of the grid-wide passes (one tile each)
constexpr int kPassStages = 4;     // radix-2 stages per grid-wide pass (16-element groups)
constexpr int kTileBlock = 512;
// The convolution length is at least kTile (any M >= 2N-1 gives the same
// cyclic convolution result), so every signal is whole tiles and a loop of up
// to 1024 points is one LDS tile pass.

struct BsLoop {
    int loop;         // index into the loop table
    int n;            // points
    int logMmax;      // log2 of the convolution length for the spanning length 2n-2
    int chunk0;       // first point chunk
    int tile0;        // first tile of its 2^logMmax-element signals
    int open;         // an open line: detrended, its ends kept
    long long xs;     // first double of the unwrapped coordinates, 3 x n
    long long chirp;  // first double2 of c_j, j < N
    long long bhat;   // first double2 of b, then DIF(b): 2^logMmax
    long long work;   // first double2 of the two coordinate signals (x + i y, z), 2 x 2^logMmax
};

struct BsState {
    int span;         // the loop spans the box: detrended and odd-extended
    int N, logM;      // transform length, log2 of the convolution length
    int kr, ki;       // kept band (kept_band)
    int filt;         // 0: nothing is zeroed and the coordinates pass through
    double coef[3], x0[3];
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// e^{sign pi i (j^2 mod 2N)/N}: exact integer argument reduction
__device__ __forceinline__ double2 chirp_at(long long j, int N, double sign)
{
    const long long q = (j * j) % (2LL * N);
    double sn, cs;
    sincospi(sign * (double)q / N, &sn, &cs);
    return make_double2(cs, sn);
}

// keep the surviving halfcomplex entries (mirrored to k > N/2) and conjugate
__device__ __forceinline__ double2 filter_conj(double2 v, int k, int N, int kr, int ki)
{
    int kk = k <= N - k ? k : N - k;
    bool keep_r = kk < kr;
    bool keep_i = kk >= 1 && kk < ki;
    return conj2(make_double2(keep_r ? v.x : 0.0, keep_i ? v.y : 0.0));
}

// u_j of coordinate c: the unwrapped loop x, detrended and odd-extended to
// length 2n-2 when it spans the box (src/vortex_smooth.c:248-279)
__device__ __forceinline__ double bs_coord(const double *x, int n, const BsState &s, int c, int j)
{
    if (!s.span) return x[j];
    const int jj = j < n ? j : 2 * n - 2 - j;
    const double v = (jj == 0 || jj == n - 1) ? 0.0 : x[jj] - s.coef[c] * jj - s.x0[c];
    return j < n ? v : -v;
}

// signal sig of a loop: 0 the packed x + i y, 1 z, 3 the b sequence
__device__ __forceinline__ double2 *bs_signal(double2 *arena, const BsLoop &B, int sig)
{
    return arena + (sig == 3 ? B.bhat : B.work + ((long long)sig << B.logMmax));
}

// Inclusive scan of three ints per thread over the block (a multiple of 64
// threads); tot receives the block totals.  sw: one row per wave.
__device__ void block_scan3(int v[3], int (*sw)[3], int tot[3])
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int off = 1; off < 64; off <<= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int u = __shfl_up(v[a], off, 64);
            if (lane >= off) v[a] += u;
        }
    if (lane == 63)
#pragma unroll
        for (int a = 0; a < 3; ++a) sw[wv][a] = v[a];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int pre = 0, t = 0;
        for (int k = 0; k < nw; ++k) {
            const int s = sw[k][a];
            pre += k < wv ? s : 0;
            t += s;
        }
        v[a] += pre;
        tot[a] = t;
    }
    __syncthreads();
}

// coordinates of point j of a loop and the periods of the link into it
__device__ __forceinline__ void point_wraps(const TangleParams &P, const int *slots, int j, int n,
                                            const double *const src[3], int w[3], double cur[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w[a] = 0;
        cur[a] = 0;
    }
    if (j >= n) return;
    const int s = slots[j];
    const int s0 = j > 0 ? slots[j - 1] : s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cur[a] = src[a][s];
        if (j > 0 && P.periodic[a]) w[a] = link_wraps(P, a, cur[a] - src[a][s0]);
    }
}

// per chunk of kChunk points: the sum of its links' periods per coordinate
__global__ __launch_bounds__(kChunk) void k_bs_wraps(
    const BsLoop *__restrict__ L, const int *__restrict__ chunk_loop,
    const int *__restrict__ loop_off, const int *__restrict__ loop_slots,
    const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
    const TangleParams *__restrict__ Pp, int *__restrict__ csum)
{
    const BsLoop B = L[chunk_loop[blockIdx.x]];
    const double *src[3] = {X, Y, Z};
    const int j = (blockIdx.x - B.chunk0) * kChunk + threadIdx.x;
    int w[3], tot[3];
    double cur[3];
    point_wraps(*Pp, loop_slots + loop_off[B.loop], j, B.n, src, w, cur);
    __shared__ int sw[kChunk / 64][3];
    block_scan3(w, sw, tot);
    if (threadIdx.x == 0) {
        csum[3 * blockIdx.x] = tot[0];
        csum[3 * blockIdx.x + 1] = tot[1];
        csum[3 * blockIdx.x + 2] = tot[2];
    }
}

// exclusive prefix of the chunk sums along each loop, one block per loop
__global__ __launch_bounds__(kChunk) void k_bs_carry(const BsLoop *__restrict__ L,
                                                     const int *__restrict__ csum,
                                                     int *__restrict__ carry)
{
    const BsLoop B = L[blockIdx.x];
    const int nc = (B.n + kChunk - 1) / kChunk;
    __shared__ int sw[kChunk / 64][3];
    int run[3] = {0, 0, 0};
    for (int base = 0; base < nc; base += kChunk) {
        const int k = base + threadIdx.x;
        int v[3], own[3], tot[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = own[a] = k < nc ? csum[3 * (B.chunk0 + k) + a] : 0;
        block_scan3(v, sw, tot);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (k < nc) carry[3 * (B.chunk0 + k) + a] = run[a] + v[a] - own[a];
            run[a] += tot[a];
        }
    }
}

// x'_j = x_j + (periods up to j) L: the unwrapped loop (block_unwrap)
__global__ __launch_bounds__(kChunk) void k_bs_unwrap(
    const BsLoop *__restrict__ L, const int *__restrict__ chunk_loop,
    const int *__restrict__ loop_off, const int *__restrict__ loop_slots,
    const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
    const TangleParams *__restrict__ Pp, const int *__restrict__ carry, double *__restrict__ arena)
{
    const BsLoop B = L[chunk_loop[blockIdx.x]];
    const TangleParams &P = *Pp;
    const double *src[3] = {X, Y, Z};
    const int j = (blockIdx.x - B.chunk0) * kChunk + threadIdx.x;
    int w[3], tot[3];
    double cur[3];
    point_wraps(P, loop_slots + loop_off[B.loop], j, B.n, src, w, cur);
    __shared__ int sw[kChunk / 64][3];
    block_scan3(w, sw, tot);
    if (j >= B.n) return;
    double *xs = arena + B.xs;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        xs[(size_t)a * B.n + j] = cur[a] + (double)(carry[3 * blockIdx.x + a] + w[a]) * P.L[a];
}

// span, detrend line, transform length and band of each loop
// (loop_spans, block_detrend, kept_band)
__global__ void k_bs_setup(int nb, const BsLoop *__restrict__ L, const double *__restrict__ arena,
                           const TangleParams *__restrict__ Pp, double fcut,
                           BsState *__restrict__ S)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const BsLoop B = L[b];
    const TangleParams &P = *Pp;
    const int n = B.n;
    const double *xs = arena + B.xs;
    double e[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) e[a] = xs[(size_t)a * n] - xs[(size_t)a * n + n - 1];
    BsState s;
    s.span = B.open || sqrt(dot3(e[0], e[1], e[2], e[0], e[1], e[2])) > (P.hi[0] - P.lo[0]) / 2;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.coef[a] = s.span ? (xs[(size_t)a * n + n - 1] - xs[(size_t)a * n]) / (n - 1) : 0.0;
        s.x0[a] = s.span ? xs[(size_t)a * n] : 0.0;
    }
    s.N = s.span ? 2 * n - 2 : n;
    s.kr = s.ki = 0;
    s.filt = kept_band(s.N, fcut, s.kr, s.ki);
    int logM = kTileLog;
    while ((1 << logM) < 2 * s.N - 1) ++logM;
    s.logM = logM;
    S[b] = s;
}

// tw[k] = e^{-2 pi i k/2^logT}, k < 2^(logT-1)
__global__ void k_bs_twiddles(int logT, double2 *__restrict__ tw)
{
    const int half = 1 << (logT - 1);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < half; k += gridDim.x * blockDim.x) {
        double sn, cs;
        sincospi(-(double)k / half, &sn, &cs);
        tw[k] = make_double2(cs, sn);
    }
}

// c_j (j < N), b_m (m < M) and the first convolution operands c_j u_j of the
// coordinates (x + i y, and z), zero padded to M; one tile per block
__global__ __launch_bounds__(256) void k_bs_tables(const BsLoop *__restrict__ L,
                                                   const BsState *__restrict__ S,
                                                   const int *__restrict__ tile_loop,
                                                   double2 *__restrict__ arena)
{
    const int b = tile_loop[blockIdx.x];
    const BsState s = S[b];
    const BsLoop B = L[b];
    const int M = 1 << s.logM;
    const int m0 = (blockIdx.x - B.tile0) << kTileLog;
    if (!s.filt || m0 >= M) return;
    const int n = B.n, N = s.N;
    const double *xs = (const double *)arena + B.xs;
    double2 *chirp = arena + B.chirp;
    double2 *bq = bs_signal(arena, B, 3);
    for (int m = m0 + threadIdx.x; m < m0 + kTile; m += blockDim.x) {
        double2 cv = make_double2(0, 0), bv = make_double2(0, 0);
        if (m < N) {
            cv = chirp_at(m, N, -1.0);
            chirp[m] = cv;
            bv = conj2(cv);
        } else if (m > M - N) {
            bv = chirp_at(M - m, N, 1.0);
        }
        bq[m] = bv;
        // x + i y in signal 0 and z in signal 1: the transforms are linear,
        // so one complex transform carries two real coordinates
        double2 v0 = make_double2(0, 0), v1 = make_double2(0, 0);
        if (m < N) {
            const double ux = bs_coord(xs, n, s, 0, m), uy = bs_coord(xs + n, n, s, 1, m);
            const double uz = bs_coord(xs + 2 * (size_t)n, n, s, 2, m);
            v0 = cmul(make_double2(ux, uy), cv);
            v1 = make_double2(uz * cv.x, uz * cv.y);
        }
        bs_signal(arena, B, 0)[m] = v0;
        bs_signal(arena, B, 1)[m] = v1;
    }
}

// One grid-wide pass over the global stages (span >= kTile) of every signal:
// stages 4*pass .. 4*pass+3, or the rest when a signal has fewer.  Forward
// (DIF): stage s has span M/2^(s+1); inverse (DIT): span kTile*2^s.  The
// twiddle of a span-h butterfly at position p is
// e^{-+2 pi i p/(2h)} = tw[p << (logT-1-log2 h)].
template <bool INV>
__global__ __launch_bounds__(kPassBlock) void k_bs_pass(
    const BsLoop *__restrict__ L, const BsState *__restrict__ S, const int *__restrict__ tile_loop,
    double2 *__restrict__ arena, const double2 *__restrict__ tw, int logT, int sig0, int pass)
{
    // up to four radix-2 stages of span >= kTile per pass: the elements
    // i0 + j hs (j < 2^ns) of a group stay in registers through them.
    // Forward (decimation in frequency): spans h, h/2, ..., from the
    // largest; inverse (decimation in time): spans hs, 2 hs, ..., from
    // kTile.  Each stage is the radix-2 step of the LDS tiles, element by
    // element, with the twiddle of the element's position in its span.
    const int b = tile_loop[blockIdx.x];
    const int logM = S[b].logM;
    if (!S[b].filt) return;
    const BsLoop B = L[b];
    const int tl = blockIdx.x - B.tile0;
    const int G = logM - kTileLog;
    const int s0 = kPassStages * pass;
    if (s0 >= G || (tl << kTileLog) >= (1 << logM)) return;
    const int ns = min(kPassStages, G - s0);
    const int R = 1 << ns;
    double2 *a = bs_signal(arena, B, sig0 + blockIdx.y);
    // log2 of the smallest span of the pass (the element stride)
    const int ls = INV ? kTileLog + s0 : logM - s0 - ns;
    const int hs = 1 << ls;
    for (int k = 0; k < (kTile >> ns) / kPassBlock; ++k) {
        const int q = (tl << (kTileLog - ns)) + k * kPassBlock + threadIdx.x;
        const int p = q & (hs - 1);
        const int i0 = ((q >> ls) << (ls + ns)) + p;
        double2 e[1 << kPassStages];
#pragma unroll
        for (int j = 0; j < (1 << kPassStages); ++j)
            if (j < R) e[j] = a[i0 + j * hs];
        if (!INV) {
            // stage t of the pass pairs elements 2^u apart, u = ns - 1 - t
            // (u runs down, so the distances stay compile-time constants)
#pragma unroll
            for (int u = kPassStages - 1; u >= 0; --u) {
                if (u < ns) {
                    const int D = 1 << u;      // pair distance in the group
                    const int lspan = ls + u;  // log2 of this stage's span
#pragma unroll
                    for (int j = 0; j < (1 << kPassStages); ++j) {
                        if (j < R && !(j & D)) {
                            const double2 w = tw[(p + (j & (D - 1)) * hs) << (logT - 1 - lspan)];
                            const double2 u = e[j], v = e[j + D];
                            e[j] = cadd(u, v);
                            e[j + D] = cmul(csub(u, v), w);
                        }
                    }
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < kPassStages; ++t) {
                if (t < ns) {
                    const int D = 1 << t;
                    const int lspan = ls + t;
#pragma unroll
                    for (int j = 0; j < (1 << kPassStages); ++j) {
                        if (j < R && !(j & D)) {
                            const double2 w =
                                conj2(tw[(p + (j & (D - 1)) * hs) << (logT - 1 - lspan)]);
                            const double2 u = e[j];
                            const double2 v = cmul(w, e[j + D]);
                            e[j] = cadd(u, v);
                            e[j + D] = csub(u, v);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < (1 << kPassStages); ++j)
            if (j < R) a[i0 + j * hs] = e[j];
    }
}

// The stages of span < kTile of one tile, in LDS: forward only (the b
// sequence), or forward, the product with DIF(b) and the inverse stages
// of span < kTile (the coordinate signals).
template <bool CONV>
__global__ __launch_bounds__(kTileBlock) void k_bs_tile(
    const BsLoop *__restrict__ L, const BsState *__restrict__ S, const int *__restrict__ tile_loop,
    double2 *__restrict__ arena, const double2 *__restrict__ tw, int logT)
{
    const int b = tile_loop[blockIdx.x];
    const int logM = S[b].logM;
    if (!S[b].filt) return;
    const BsLoop B = L[b];
    const int tl = blockIdx.x - B.tile0;
    if ((tl << kTileLog) >= (1 << logM)) return;
    double2 *a = bs_signal(arena, B, CONV ? (int)blockIdx.y : 3) + ((size_t)tl << kTileLog);
    __shared__ double2 sh[kTile];
    for (int i = threadIdx.x; i < kTile; i += kTileBlock) sh[i] = a[i];
    __syncthreads();
    // rounds of three radix-2 stages, one 8-element group per thread held in
    // registers (element stride hs = the round's smallest span), the stage
    // order and twiddles of the radix-2 sweeps: forward spans 2048 .. 1,
    // inverse spans 1 .. 2048
    static_assert(kTileLog % 3 == 0 && kTileBlock == kTile / 8, "8-element groups");
    const int q = threadIdx.x;
    for (int r = kTileLog / 3 - 1; r >= 0; --r) {
        const int ls = 3 * r, hs = 1 << ls;
        const int p = q & (hs - 1), i0 = ((q >> ls) << (ls + 3)) + p;
        double2 e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = sh[i0 + j * hs];
#pragma unroll
        for (int u = 2; u >= 0; --u) {
            const int D = 1 << u;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (!(j & D)) {
                    const double2 w = tw[(p + (j & (D - 1)) * hs) << (logT - 1 - (ls + u))];
                    const double2 x = e[j], y = e[j + D];
                    e[j] = cadd(x, y);
                    e[j + D] = cmul(csub(x, y), w);
                }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) sh[i0 + j * hs] = e[j];
        __syncthreads();
    }
    if (CONV) {
        const double2 *bh = bs_signal(arena, B, 3) + ((size_t)tl << kTileLog);
        for (int i = threadIdx.x; i < kTile; i += kTileBlock) sh[i] = cmul(sh[i], bh[i]);
        __syncthreads();
        for (int r = 0; r < kTileLog / 3; ++r) {
            const int ls = 3 * r, hs = 1 << ls;
            const int p = q & (hs - 1), i0 = ((q >> ls) << (ls + 3)) + p;
            double2 e[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) e[j] = sh[i0 + j * hs];
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const int D = 1 << u;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (!(j & D)) {
                        const double2 w =
                            conj2(tw[(p + (j & (D - 1)) * hs) << (logT - 1 - (ls + u))]);
                        const double2 x = e[j], y = cmul(w, e[j + D]);
                        e[j] = cadd(x, y);
                        e[j + D] = csub(x, y);
                    }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) sh[i0 + j * hs] = e[j];
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < kTile; i += kTileBlock) a[i] = sh[i];
}

// Between the two transforms: U = c . conv / M, filtered and conjugated, and
// times c again as the second convolution operand (zero padded).
__global__ __launch_bounds__(256) void k_bs_mid(const BsLoop *__restrict__ L,
                                                const BsState *__restrict__ S,
                                                const int *__restrict__ tile_loop,
                                                double2 *__restrict__ arena)
{
    const int b = tile_loop[blockIdx.x];
    const BsState s = S[b];
    const BsLoop B = L[b];
    const int M = 1 << s.logM;
    const int m0 = (blockIdx.x - B.tile0) << kTileLog;
    if (!s.filt || m0 >= M) return;
    double2 *a = bs_signal(arena, B, blockIdx.y);
    const double2 *chirp = arena + B.chirp;
    const double inv = 1.0 / M;
    const int N = s.N;
    for (int m = m0 + threadIdx.x; m < m0 + kTile; m += blockDim.x) {
        if (m >= N) {
            a[m] = make_double2(0, 0);
            continue;
        }
        const double2 cm = chirp[m];
        if (blockIdx.y == 1) {  // z alone
            const double2 u = cmul(a[m], cm);
            a[m] = cmul(filter_conj(make_double2(u.x * inv, u.y * inv), m, N, s.kr, s.ki), cm);
            continue;
        }
        // x + i y: Z_m = X_m + i Y_m with X, Y the spectra of real signals,
        // X_m = (Z_m + conj Z_n) / 2 and Y_m = (Z_m - conj Z_n) / 2i for
        // n = N - m.  The filter (real and imaginary half-complex parts kept
        // separately) acts on X and Y; it maps a Hermitian spectrum to a
        // Hermitian one, so X' + i Y' is again the transform of x' + i y'.
        // The thread of the lower index of each pair (m, n) writes both.
        const int n = m == 0 ? 0 : N - m;
        if (n < m) continue;
        const double2 cn = chirp[n];
        const double2 zm = cmul(a[m], cm), zn = cmul(a[n], cn);
        const double2 X = make_double2(0.5 * inv * (zm.x + zn.x), 0.5 * inv * (zm.y - zn.y));
        const double2 Y = make_double2(0.5 * inv * (zm.y + zn.y), 0.5 * inv * (zn.x - zm.x));
        const double2 Xf = conj2(filter_conj(X, m, N, s.kr, s.ki));
        const double2 Yf = conj2(filter_conj(Y, m, N, s.kr, s.ki));
        // Z'_m = X' + i Y', Z'_n = conj X' + i conj Y'; the second transform
        // takes their conjugates times the chirp
        a[m] = cmul(make_double2(Xf.x - Yf.y, -(Xf.y + Yf.x)), cm);
        if (n != m) a[n] = cmul(make_double2(Xf.x + Yf.y, Xf.y - Yf.x), cn);
    }
}

// x_j, y_j, z_j from c_j conv_j / (M N), the detrend line added back, written along the
// loop (src/vortex_smooth.c:286-309)
__global__ __launch_bounds__(kChunk) void k_bs_finish(
    const BsLoop *__restrict__ L, const BsState *__restrict__ S, const int *__restrict__ chunk_loop,
    const int *__restrict__ loop_off, const int *__restrict__ loop_slots,
    double2 *__restrict__ arena, double *__restrict__ X, double *__restrict__ Y,
    double *__restrict__ Z)
{
    const int b = chunk_loop[blockIdx.x];
    const BsLoop B = L[b];
    const int j = (blockIdx.x - B.chunk0) * kChunk + threadIdx.x;
    if (j >= B.n) return;
    if (B.open && (j == 0 || j == B.n - 1)) return;  // an open line keeps its ends
    const BsState s = S[b];
    const double *xs = (const double *)arena + B.xs;
    const double inv = 1.0 / (1 << s.logM);
    double v[3];
    if (s.filt) {
        // the second transform of conj(X' + i Y') is N (x' - i y')
        const double2 cj = arena[B.chirp + j];
        const double2 u0 = cmul(bs_signal(arena, B, 0)[j], cj);
        const double2 u1 = cmul(bs_signal(arena, B, 1)[j], cj);
        v[0] = u0.x * inv / s.N;
        v[1] = -u0.y * inv / s.N;
        v[2] = u1.x * inv / s.N;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = bs_coord(xs + (size_t)c * B.n, B.n, s, c, j);
    }
    if (s.span)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = v[c] + s.coef[c] * j + s.x0[c];
    const int m = loop_slots[loop_off[B.loop] + j];
    X[m] = v[0];
    Y[m] = v[1];
    Z[m] = v[2];
}

static inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// Smooth the loops of the loop table listed in `big` (all longer than
// kSmallLoop) with the batched transform.  hmeta (the context's) holds the
// uploaded tables until the next call.
static int smooth_big(Ctx &c, const std::vector<int> &big, std::vector<char> &hmeta)
{
    const int nb = (int)big.size();
    std::vector<BsLoop> L(nb);
    std::vector<int> chunk_loop, tile_loop;
    long long arena = 0;  // double2
    int logMall = kTileLog;
    auto al16 = [](long long v) { return (v + 15) & ~15LL; };
    for (int i = 0; i < nb; ++i) {
        const int lp = big[i];
        const int n = c.h_loop_off[lp + 1] - c.h_loop_off[lp];
        const long long Nmax = 2LL * n - 2;
        int logM = kTileLog;
        while ((1LL << logM) < 2 * Nmax - 1) ++logM;
        if (logM > 30) return set_error(OVG_ELIMIT, "a loop of %d points is too long to smooth", n);
        BsLoop &B = L[i];
        B.loop = lp;
        B.n = n;
        B.logMmax = logM;
        B.open = c.h_loop_open.empty() ? 0 : c.h_loop_open[lp];
        B.chunk0 = (int)chunk_loop.size();
        B.tile0 = (int)tile_loop.size();
        chunk_loop.insert(chunk_loop.end(), (n + kChunk - 1) / kChunk, i);
        tile_loop.insert(tile_loop.end(), (size_t)1 << (logM - kTileLog), i);
        B.xs = 2 * arena;
        arena = al16(arena + (3LL * n + 1) / 2);
        B.chirp = arena;
        arena = al16(arena + Nmax);
        B.bhat = arena;
        arena = al16(arena + (1LL << logM));
        B.work = arena;
        arena = al16(arena + (2LL << logM));
        logMall = std::max(logMall, logM);
    }
    if (tile_loop.size() > (size_t)INT32_MAX / 2)
        return set_error(OVG_ELIMIT, "too many transform tiles to smooth");
    const long long nch = (long long)chunk_loop.size(), nt = (long long)tile_loop.size();
    auto al256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t offC = al256(sizeof(BsLoop) * nb), offT = offC + al256(sizeof(int) * nch);
    const size_t meta = offT + sizeof(int) * nt;
    const size_t offS = 0, offW = al256(sizeof(BsState) * nb), offK = offW + al256(12 * nch);
    const size_t state = offK + 12 * nch;
    OVG_TRY(c.sm_big.ensure(meta));
    OVG_TRY(c.sm_gstate.ensure(state));
    OVG_TRY(c.sm_fft.ensure(sizeof(double2) * (size_t)arena));
    if (logMall > c.sm_tw_log) {
        OVG_TRY(c.sm_tw.ensure(sizeof(double2) << (logMall - 1)));
        c.sm_tw_log = logMall;
        hipLaunchKernelGGL(k_bs_twiddles, dim3(nblocks(1LL << (logMall - 1), 256)), dim3(256), 0,
                           c.stream, c.sm_tw_log, c.sm_tw.as<double2>());
    }
    if (c.sm_hmeta_done) OVG_HIP(hipEventSynchronize(c.sm_hmeta_done));  // the last upload
    hmeta.assign(meta, 0);
    std::memcpy(hmeta.data(), L.data(), sizeof(BsLoop) * nb);
    std::memcpy(hmeta.data() + offC, chunk_loop.data(), sizeof(int) * nch);
    std::memcpy(hmeta.data() + offT, tile_loop.data(), sizeof(int) * nt);
    OVG_HIP(hipMemcpyAsync(c.sm_big.p, hmeta.data(), meta, hipMemcpyHostToDevice, c.stream));
    if (!c.sm_hmeta_done)
        OVG_HIP(hipEventCreateWithFlags(&c.sm_hmeta_done, hipEventDisableTiming));
    OVG_HIP(hipEventRecord(c.sm_hmeta_done, c.stream));

    char *mb = c.sm_big.as<char>(), *sb = c.sm_gstate.as<char>();
    const BsLoop *dL = (const BsLoop *)mb;
    const int *dC = (const int *)(mb + offC), *dT = (const int *)(mb + offT);
    BsState *dS = (BsState *)(sb + offS);
    int *csum = (int *)(sb + offW), *carry = (int *)(sb + offK);
    double2 *ar = c.sm_fft.as<double2>();
    const double2 *tw = c.sm_tw.as<double2>();
    const int logT = c.sm_tw_log;
    const int *loff = c.loop_off.as<int>(), *lsl = c.loop_slots.as<int>();
    double *X = c.x[0].as<double>(), *Y = c.x[1].as<double>(), *Z = c.x[2].as<double>();
    const int passes = (logMall - kTileLog + kPassStages - 1) / kPassStages;
    const dim3 one((unsigned)nt, 1), two((unsigned)nt, 2);

    hipLaunchKernelGGL(k_bs_wraps, dim3((unsigned)nch), dim3(kChunk), 0, c.stream, dL, dC, loff,
                       lsl, X, Y, Z, dparams(c), csum);
    hipLaunchKernelGGL(k_bs_carry, dim3(nb), dim3(kChunk), 0, c.stream, dL, csum, carry);
    hipLaunchKernelGGL(k_bs_unwrap, dim3((unsigned)nch), dim3(kChunk), 0, c.stream, dL, dC, loff,
                       lsl, X, Y, Z, dparams(c), carry, (double *)ar);
    hipLaunchKernelGGL(k_bs_setup, dim3(nblocks(nb, 64)), dim3(64), 0, c.stream, nb, dL,
                       (const double *)ar, dparams(c), c.params.freq_to_cutoff, dS);
    hipLaunchKernelGGL(k_bs_tables, one, dim3(256), 0, c.stream, dL, dS, dT, ar);
    for (int q = 0; q < passes; ++q)
        hipLaunchKernelGGL((k_bs_pass<false>), one, dim3(kPassBlock), 0, c.stream, dL, dS, dT, ar,
                           tw, logT, 3, q);
    hipLaunchKernelGGL((k_bs_tile<false>), one, dim3(kTileBlock), 0, c.stream, dL, dS, dT, ar, tw,
                       logT);
    for (int rep = 0; rep < 2; ++rep) {
        for (int q = 0; q < passes; ++q)
            hipLaunchKernelGGL((k_bs_pass<false>), two, dim3(kPassBlock), 0, c.stream, dL, dS,
                               dT, ar, tw, logT, 0, q);
        hipLaunchKernelGGL((k_bs_tile<true>), two, dim3(kTileBlock), 0, c.stream, dL, dS, dT,
                           ar, tw, logT);
        for (int q = 0; q < passes; ++q)
            hipLaunchKernelGGL((k_bs_pass<true>), two, dim3(kPassBlock), 0, c.stream, dL, dS,
                               dT, ar, tw, logT, 0, q);
        if (rep == 0)
            hipLaunchKernelGGL(k_bs_mid, two, dim3(256), 0, c.stream, dL, dS, dT, ar);
    }
    hipLaunchKernelGGL(k_bs_finish, dim3((unsigned)nch), dim3(kChunk), 0, c.stream, dL, dS, dC,
                       loff, lsl, ar, X, Y, Z);
    return OVG_OK;
}

// ---- loop table ------------------------------------------------------------
// Loops in order of their smallest slot, each listed forward from it
// (src/vortex_smooth.c:192-242).  On the device: min-label pointer jumping
// over the forward links gives every slot the smallest slot of its loop (a
// window of 2^r links per round: ceil(log2 n_live) rounds cover any loop),
// list ranking (Wyllie) on each loop cut open at that slot gives every
// slot's distance d to it, so its position is (len - d) mod len; heads are
// numbered by an exclusive scan in slot order and every slot is scattered
// to its place.  Open chains (a live forward link < 0) take the host walk.

__global__ void k_lt_init(int n, const int *__restrict__ fwd, const int *__restrict__ status,
                          int *__restrict__ lab, int *__restrict__ nx, int *__restrict__ open)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool live = status[i] != OVG_EMPTY;  // one status word per slot
    const int f = fwd[i];
    // open chains, and links into EMPTY or missing slots, take the host walk
    if (live && (f < 0 || f >= n || status[f] == OVG_EMPTY)) *open = 1;
    lab[i] = live ? i : INT_MAX;
    nx[i] = (live && f >= 0 && f < n) ? f : i;
}

__global__ void k_lt_minjump(int n, const int *__restrict__ lab, const int *__restrict__ nx,
                             int *__restrict__ lab2, int *__restrict__ nx2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = nx[i];
    lab2[i] = min(lab[i], lab[j]);
    nx2[i] = nx[j];
}

// the head (smallest slot) ends its list; dead slots are lists of their own
__global__ void k_lt_rank_init(int n, const int *__restrict__ fwd, const int *__restrict__ lab,
                               int *__restrict__ nx, int *__restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool end = lab[i] == INT_MAX || lab[i] == i || fwd[i] < 0 || fwd[i] >= n;
    nx[i] = end ? i : fwd[i];
    d[i] = end ? 0 : 1;
}

__global__ void k_lt_rankjump(int n, const int *__restrict__ nx, const int *__restrict__ d,
                              int *__restrict__ nx2, int *__restrict__ d2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = nx[i];
    d2[i] = d[i] + d[j];
    nx2[i] = nx[j];
}

// head flags and loop lengths at the heads (d of the head's successor + 1)
__global__ void k_lt_heads(int n, const int *__restrict__ fwd, const int *__restrict__ lab,
                           const int *__restrict__ d, int *__restrict__ flag, int *__restrict__ len)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool head = lab[i] == i;
    flag[i] = head;
    const int f = fwd[i];
    len[i] = head ? (f >= 0 && f < n ? d[f] + 1 : 1) : 0;  // open chains: host walk
}

// hidx: loop number of each head, hoff: first entry of each head's loop
// (exclusive scans of flag and len); position (len - d) mod len
__global__ void k_lt_scatter(int n, const int *__restrict__ lab, const int *__restrict__ d,
                             const int *__restrict__ len, const int *__restrict__ hidx,
                             const int *__restrict__ hoff, int *__restrict__ loop_off,
                             int *__restrict__ slots)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) loop_off[hidx[n]] = hoff[n];
    if (i >= n || lab[i] == INT_MAX) return;
    const int m = lab[i];
    if (m == i) loop_off[hidx[i]] = hoff[i];
    const int pos = m == i ? 0 : len[m] - d[i];
    if (pos >= 0 && pos < len[m]) slots[hoff[m] + pos] = i;  // always, for consistent links
}

__global__ void k_fill_zero(int n, int *a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// Loop table from the host link mirror (chains with open ends): loops in
// the order of their smallest slot i; a closed loop listed forward from i,
// an open line forward from its start, the point is_vortex_open reaches by
// the reverse links from i (src/vortex_smooth.c:24-39,198-205).
static int build_loops_host(Ctx &c)
{
    int n = c.n;
    std::vector<int> off, slots, open;
    std::vector<char> seen(n, 0);
    off.push_back(0);
    for (int i = 0; i < n; ++i) {
        if (seen[i] || c.h_status[2 * i] == OVG_EMPTY) continue;
        int s = i, guard = 0;
        int r = c.h_conn[2 * s + 1];
        while (r != i && r >= 0 && r < n && ++guard <= n) {
            s = r;
            r = c.h_conn[2 * s + 1];
        }
        const bool is_open = r < 0 || r >= n;
        const int start = is_open ? s : i;
        int k = start;
        do {
            seen[k] = 1;
            slots.push_back(k);
            k = c.h_conn[2 * k];
        } while (k != start && k >= 0 && k < n && !seen[k]);
        off.push_back((int)slots.size());
        open.push_back(is_open ? 1 : 0);
    }
    c.n_loops = (int)off.size() - 1;
    c.h_loop_off = off;
    c.h_loop_open = open;
    OVG_TRY(c.loop_off.ensure(sizeof(int) * off.size()));
    OVG_TRY(c.loop_slots.ensure(sizeof(int) * (slots.size() + 1)));
    OVG_TRY(c.loop_open.ensure(sizeof(int) * (open.size() + 1)));
    OVG_HIP(hipMemcpy(c.loop_off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    if (!slots.empty())
        OVG_HIP(hipMemcpy(c.loop_slots.p, slots.data(), sizeof(int) * slots.size(),
                          hipMemcpyHostToDevice));
    if (!open.empty())
        OVG_HIP(hipMemcpy(c.loop_open.p, open.data(), sizeof(int) * open.size(),
                          hipMemcpyHostToDevice));
    c.loops_version = c.topo_version;
    return OVG_OK;
}

static int build_loops(Ctx &c)
{
    if (c.loops_version == c.topo_version) return OVG_OK;
    const int n = c.n;
    if (n <= 0) return build_loops_host(c);
    // lab, nx, lab2, nx2, len, flag (n + 1), len scan (n + 1), open flag
    const size_t stride = ((size_t)n + 64) & ~(size_t)63;
    OVG_TRY(c.lt_buf.ensure(sizeof(int) * (7 * stride + 64)));
    int *A = c.lt_buf.as<int>();
    int *lab = A, *nx = A + stride, *lab2 = A + 2 * stride, *nx2 = A + 3 * stride;
    int *len = A + 4 * stride, *flag = A + 5 * stride, *hoff = A + 6 * stride;
    int *open = A + 7 * stride;
    const unsigned g = nblocks(n, 256);
    OVG_HIP(hipMemsetAsync(open, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(k_lt_init, dim3(g), dim3(256), 0, c.stream, n, c.fwd.as<int>(),
                       c.status.as<int>(), lab, nx, open);
    int rounds = 0;
    while ((1LL << rounds) < (long long)std::max(n, 2)) ++rounds;
    for (int r = 0; r < rounds; ++r) {
        hipLaunchKernelGGL(k_lt_minjump, dim3(g), dim3(256), 0, c.stream, n, lab, nx, lab2, nx2);
        std::swap(lab, lab2);
        std::swap(nx, nx2);
    }
    // lab is final; nx, lab2, nx2 are free: ranking in nx/lab2 <-> nx2/len
    int *d = lab2, *d2 = len;
    hipLaunchKernelGGL(k_lt_rank_init, dim3(g), dim3(256), 0, c.stream, n, c.fwd.as<int>(), lab,
                       nx, d);
    for (int r = 0; r < rounds; ++r) {
        hipLaunchKernelGGL(k_lt_rankjump, dim3(g), dim3(256), 0, c.stream, n, nx, d, nx2, d2);
        std::swap(nx, nx2);
        std::swap(d, d2);
    }
    int *L = d2;  // the free one of lab2 / len
    hipLaunchKernelGGL(k_lt_heads, dim3(g), dim3(256), 0, c.stream, n, c.fwd.as<int>(), lab, d,
                       flag, L);
    OVG_TRY(launch_exclusive_scan(c, flag, flag, n, flag + n, c.tree.scan_tmp));
    OVG_TRY(launch_exclusive_scan(c, L, hoff, n, hoff + n, c.tree.scan_tmp));
    int hv[3] = {0, 0, 0};
    OVG_HIP(hipMemcpyAsync(&hv[0], flag + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipMemcpyAsync(&hv[1], hoff + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipMemcpyAsync(&hv[2], open, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    if (hv[2]) return build_loops_host(c);
    const int n_loops = hv[0];
    OVG_TRY(c.loop_off.ensure(sizeof(int) * ((size_t)n_loops + 1)));
    OVG_TRY(c.loop_slots.ensure(sizeof(int) * ((size_t)hv[1] + 1)));
    hipLaunchKernelGGL(k_lt_scatter, dim3(g), dim3(256), 0, c.stream, n, lab, d, L, flag, hoff,
                       c.loop_off.as<int>(), c.loop_slots.as<int>());
    c.n_loops = n_loops;
    c.h_loop_open.clear();  // closed loops only
    OVG_TRY(c.loop_open.ensure(sizeof(int) * ((size_t)n_loops + 1)));
    hipLaunchKernelGGL(k_fill_zero, dim3(nblocks(n_loops + 1, 256)), dim3(256), 0, c.stream,
                       n_loops + 1, c.loop_open.as<int>());
    c.h_loop_off.resize((size_t)n_loops + 1);
    OVG_HIP(hipMemcpyAsync(c.h_loop_off.data(), c.loop_off.p, sizeof(int) * ((size_t)n_loops + 1),
                           hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    c.loops_version = c.topo_version;
    return OVG_OK;
}


int smooth_loops(Ctx &c)
{
    OVG_TRY(build_loops(c));
    if (c.n_loops == 0) return OVG_OK;
    OVG_TRY(positions_change(c));
    size_t need = sizeof(double) * ((size_t)15 * c.n_live + (size_t)6 * c.n_loops + 64);
    OVG_TRY(c.sm_buf.ensure(need));
    std::vector<int> big;
    for (int lp = 0; lp < c.n_loops; ++lp)
        if (c.h_loop_off[lp + 1] - c.h_loop_off[lp] > kSmallLoop) big.push_back(lp);
    hipLaunchKernelGGL(k_smooth, dim3(c.n_loops), dim3(kSmBlock), 0, c.stream, c.n_loops,
                       c.loop_off.as<int>(), c.loop_slots.as<int>(), c.loop_open.as<int>(),
                       c.x[0].as<double>(),
                       c.x[1].as<double>(), c.x[2].as<double>(), dparams(c),
                       c.params.freq_to_cutoff, c.sm_buf.as<double>());
    // no wait here: the tables' host copy lives in the context, and the next
    // call waits for its upload before rewriting it
    if (!big.empty()) OVG_TRY(smooth_big(c, big, c.sm_hmeta));
    return OVG_OK;
}

}  // namespace ovg
