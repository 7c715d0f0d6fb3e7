// Frequency-filter smoothing (smooth -> high_freq_cutoff -> low_pass_filter,
// src/vortex_smooth.c:134-175,183-453,459-462), one workgroup per closed
// loop.
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
#include <vector>

#include "engine_internal.h"

namespace ovg {

constexpr int kSmBlock = 256;
constexpr int kSmallLoop = 1024;     // DFT of the kept band, 256 threads per loop
constexpr int kMediumLoop = 65536;   // Bluestein FFT inside one 1024-thread workgroup
constexpr int kFftBlock = 1024;

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
                if (j > 0 && P.periodic[a]) {
                    double d = cur - src[a][slots[j - 1]];
                    int g = 0;
                    while (d > P.halfL[a] && g++ < 64) { d -= P.L[a]; --w; }
                    while (d < -P.halfL[a] && g++ < 128) { d += P.L[a]; ++w; }
                }
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
__device__ void block_finish(double *const xs[3], int n, bool span, const double *coef,
                             const double *x0, const int *slots, double *X, double *Y, double *Z)
{
    if (span) {
        for (int j = threadIdx.x; j < n; j += blockDim.x)
            for (int c = 0; c < 3; ++c) xs[c][j] = xs[c][j] + coef[c] * j + x0[c];
        __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
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
    double *__restrict__ X, double *__restrict__ Y, double *__restrict__ Z,
    const TangleParams *__restrict__ Pp, double fcut, double *__restrict__ work)
{
    int lp = blockIdx.x;
    if (lp >= n_loops) return;
    const TangleParams &P = *Pp;
    const int off = loop_off[lp];
    const int n = loop_off[lp + 1] - off;
    if (n > kSmallLoop) return;  // FFT paths
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
    if (threadIdx.x == 0) s_span = loop_spans(P, xs, n);
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
    block_finish(xs, n, span, s_coef, s_x0, slots, X, Y, Z);
}

// ---- larger loops: Bluestein transform over radix-2 FFTs --------------------
// U_k = sum_j u_j e^{-2 pi i jk/N} = c_k (a * b)_k with the chirp
// c_j = e^{-pi i j^2/N}, a_j = u_j c_j, b_m = conj(c_m) (Bluestein 1970); the
// cyclic convolution of length M = 2^ceil(log2(2N-1)) uses radix-2 FFTs.
// The filter zeroes the same halfcomplex entries as the small-loop path and
// y = conj(DFT(conj(U)))/N.

struct BigLoop {
    int loop;       // index into the loop table
    int logM;       // for the spanning length 2n-2 (an upper bound otherwise)
    long long off;  // first double2 of this loop's scratch
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// in place, within the workgroup; tw[k] = e^{-2 pi i k/M}, k < M/2
__device__ void fft_radix2(double2 *a, int M, int logM, const double2 *tw, bool inverse)
{
    for (int i = threadIdx.x; i < M; i += blockDim.x) {
        int j = (int)(__brev((unsigned)i) >> (32 - logM));
        if (i < j) {
            double2 t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    __syncthreads();
    for (int st = 1; st <= logM; ++st) {
        int half = 1 << (st - 1);
        int stride = M >> st;
        for (int b = threadIdx.x; b < M / 2; b += blockDim.x) {
            int grp = b >> (st - 1), pos = b & (half - 1);
            int i = grp * 2 * half + pos, j = i + half;
            double2 w = tw[pos * stride];
            if (inverse) w = conj2(w);
            double2 t = cmul(w, a[j]);
            double2 u = a[i];
            a[i] = make_double2(u.x + t.x, u.y + t.y);
            a[j] = make_double2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }
}

__device__ void bluestein(double2 *U, int N, int M, int logM, const double2 *chirp,
                          const double2 *Bhat, const double2 *tw, double2 *work)
{
    for (int j = threadIdx.x; j < M; j += blockDim.x)
        work[j] = j < N ? cmul(U[j], chirp[j]) : make_double2(0, 0);
    __syncthreads();
    fft_radix2(work, M, logM, tw, false);
    for (int j = threadIdx.x; j < M; j += blockDim.x) work[j] = cmul(work[j], Bhat[j]);
    __syncthreads();
    fft_radix2(work, M, logM, tw, true);
    const double inv = 1.0 / M;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        double2 v = cmul(work[k], chirp[k]);
        U[k] = make_double2(v.x * inv, v.y * inv);
    }
    __syncthreads();
}

// twiddles, chirp (exact argument reduction) and the b sequence
__device__ void bluestein_tables(int N, int M, double2 *tw, double2 *chirp, double2 *Bhat,
                                 int tid, int nthreads)
{
    for (int k = tid; k < M / 2; k += nthreads) {
        double sn, cs;
        sincospi(-2.0 * (double)k / M, &sn, &cs);
        tw[k] = make_double2(cs, sn);
    }
    for (int j = tid; j < N; j += nthreads) {
        long long q = ((long long)j * j) % (2LL * N);
        double sn, cs;
        sincospi(-(double)q / N, &sn, &cs);
        chirp[j] = make_double2(cs, sn);
    }
    for (int m = tid; m < M; m += nthreads) {
        double2 v = make_double2(0, 0);
        int mm = m < N ? m : (m > M - N ? M - m : -1);
        if (mm >= 0) {
            long long q = ((long long)mm * mm) % (2LL * N);
            double sn, cs;
            sincospi((double)q / N, &sn, &cs);  // conj(chirp)
            v = make_double2(cs, sn);
        }
        Bhat[m] = v;
    }
}

// keep the surviving halfcomplex entries (mirrored to k > N/2) and conjugate
__device__ __forceinline__ double2 filter_conj(double2 v, int k, int N, int kr, int ki)
{
    int kk = k <= N - k ? k : N - k;
    bool keep_r = kk < kr;
    bool keep_i = kk >= 1 && kk < ki;
    return conj2(make_double2(keep_r ? v.x : 0.0, keep_i ? v.y : 0.0));
}

__global__ __launch_bounds__(kFftBlock) void k_smooth_fft(
    int n_big, const BigLoop *__restrict__ big, const int *__restrict__ loop_off,
    const int *__restrict__ loop_slots, double *__restrict__ X, double *__restrict__ Y,
    double *__restrict__ Z, const TangleParams *__restrict__ Pp, double fcut,
    double2 *__restrict__ scratch)
{
    const BigLoop B = big[blockIdx.x];
    const TangleParams &P = *Pp;
    const int off = loop_off[B.loop];
    const int n = loop_off[B.loop + 1] - off;
    const int *slots = loop_slots + off;
    // scratch: xs (3 x 2n doubles = 3n double2), tw (M/2), chirp (2n),
    // Bhat (M), work (M), U (2n); M sized for the spanning length
    const int Mmax = 1 << B.logM;
    double2 *base = scratch + B.off;
    double *xs[3] = {(double *)base, (double *)base + 2 * (size_t)n,
                     (double *)base + 4 * (size_t)n};
    double2 *tw = base + 3 * (size_t)n;
    double2 *chirp = tw + Mmax / 2;
    double2 *Bhat = chirp + 2 * (size_t)n;
    double2 *work = Bhat + Mmax;
    double2 *U = work + Mmax;
    __shared__ int sh[kFftBlock];
    __shared__ int s_carry, s_span;
    __shared__ double s_coef[3], s_x0[3];
    block_unwrap(P, slots, n, X, Y, Z, xs, sh, &s_carry);
    if (threadIdx.x == 0) s_span = loop_spans(P, xs, n);
    __syncthreads();
    const bool span = s_span;
    if (span) block_detrend(xs, n, s_coef, s_x0);
    const int N = span ? 2 * n - 2 : n;
    int kr, ki;
    if (kept_band(N, fcut, kr, ki)) {
        int logM = 0;
        while ((1 << logM) < 2 * N - 1) ++logM;
        const int M = 1 << logM;
        bluestein_tables(N, M, tw, chirp, Bhat, threadIdx.x, blockDim.x);
        __syncthreads();
        fft_radix2(Bhat, M, logM, tw, false);
        for (int c = 0; c < 3; ++c) {
            for (int j = threadIdx.x; j < N; j += blockDim.x) U[j] = make_double2(xs[c][j], 0.0);
            __syncthreads();
            bluestein(U, N, M, logM, chirp, Bhat, tw, work);
            for (int k = threadIdx.x; k < N; k += blockDim.x) U[k] = filter_conj(U[k], k, N, kr, ki);
            __syncthreads();
            bluestein(U, N, M, logM, chirp, Bhat, tw, work);
            for (int j = threadIdx.x; j < n; j += blockDim.x) xs[c][j] = U[j].x / N;
            __syncthreads();
        }
    }
    block_finish(xs, n, span, s_coef, s_x0, slots, X, Y, Z);
}

// ---- giant loops: the same transform with grid-wide FFT stages ------------
struct GiantState {
    int span;
    int pad;
    double coef[3], x0[3];
};

__global__ __launch_bounds__(kFftBlock) void k_giant_unwrap(
    int loop, const int *__restrict__ loop_off, const int *__restrict__ loop_slots,
    const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
    const TangleParams *__restrict__ Pp, double *__restrict__ xs0, GiantState *__restrict__ gs)
{
    const TangleParams &P = *Pp;
    const int off = loop_off[loop];
    const int n = loop_off[loop + 1] - off;
    double *xs[3] = {xs0, xs0 + 2 * (size_t)n, xs0 + 4 * (size_t)n};
    __shared__ int sh[kFftBlock];
    __shared__ int s_carry, s_span;
    __shared__ double s_coef[3], s_x0[3];
    block_unwrap(P, loop_slots + off, n, X, Y, Z, xs, sh, &s_carry);
    if (threadIdx.x == 0) s_span = loop_spans(P, xs, n);
    __syncthreads();
    if (s_span) block_detrend(xs, n, s_coef, s_x0);
    if (threadIdx.x == 0) {
        gs->span = s_span;
        for (int c = 0; c < 3; ++c) {
            gs->coef[c] = s_span ? s_coef[c] : 0;
            gs->x0[c] = s_span ? s_x0[c] : 0;
        }
    }
}

__global__ void k_g_tables(int N, int M, double2 *tw, double2 *chirp, double2 *Bhat)
{
    bluestein_tables(N, M, tw, chirp, Bhat, blockIdx.x * blockDim.x + threadIdx.x,
                     gridDim.x * blockDim.x);
}

__global__ void k_g_bitrev(double2 *a, int M, int logM)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int j = (int)(__brev((unsigned)i) >> (32 - logM));
    if (i < j) {
        double2 t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
}

__global__ void k_g_stage(double2 *a, int M, int st, const double2 *tw, int inverse)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= M / 2) return;
    int half = 1 << (st - 1);
    int grp = b >> (st - 1), pos = b & (half - 1);
    int i = grp * 2 * half + pos, j = i + half;
    double2 w = tw[pos * (M >> st)];
    if (inverse) w = conj2(w);
    double2 t = cmul(w, a[j]);
    double2 u = a[i];
    a[i] = make_double2(u.x + t.x, u.y + t.y);
    a[j] = make_double2(u.x - t.x, u.y - t.y);
}

// mode 0: work = U c (zero padded); 1: work *= Bhat; 2: U = work c / M
__global__ void k_g_point(int mode, double2 *work, double2 *U, const double2 *chirp,
                          const double2 *Bhat, int N, int M)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    if (mode == 0) {
        work[j] = j < N ? cmul(U[j], chirp[j]) : make_double2(0, 0);
    } else if (mode == 1) {
        work[j] = cmul(work[j], Bhat[j]);
    } else if (j < N) {
        double2 v = cmul(work[j], chirp[j]);
        U[j] = make_double2(v.x / M, v.y / M);
    }
}

// mode 0: U = xs (real); 1: U = conj(filter(U)); 2: xs = Re U / N (j < n)
__global__ void k_g_coord(int mode, double *xs, double2 *U, int n, int N, int kr, int ki)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 0) {
        if (j < N) U[j] = make_double2(xs[j], 0.0);
    } else if (mode == 1) {
        if (j < N) U[j] = filter_conj(U[j], j, N, kr, ki);
    } else if (j < n) {
        xs[j] = U[j].x / N;
    }
}

__global__ void k_giant_finish(int loop, const int *__restrict__ loop_off,
                               const int *__restrict__ loop_slots, double *__restrict__ xs0,
                               const GiantState *__restrict__ gs, double *__restrict__ X,
                               double *__restrict__ Y, double *__restrict__ Z)
{
    const int off = loop_off[loop];
    const int n = loop_off[loop + 1] - off;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double v[3];
    for (int c = 0; c < 3; ++c) {
        v[c] = xs0[2 * (size_t)n * c + j];
        if (gs->span) v[c] = v[c] + gs->coef[c] * j + gs->x0[c];
    }
    int m = loop_slots[off + j];
    X[m] = v[0];
    Y[m] = v[1];
    Z[m] = v[2];
}

static inline int g1(long long n) { return (int)((n + 255) / 256); }

static void fft_grid(Ctx &c, double2 *a, int M, int logM, const double2 *tw, int inverse)
{
    hipLaunchKernelGGL(k_g_bitrev, dim3(g1(M)), dim3(256), 0, c.stream, a, M, logM);
    for (int st = 1; st <= logM; ++st)
        hipLaunchKernelGGL(k_g_stage, dim3(g1(M / 2)), dim3(256), 0, c.stream, a, M, st, tw,
                           inverse);
}

static void bluestein_grid(Ctx &c, double2 *U, int N, int M, int logM, const double2 *chirp,
                           const double2 *Bhat, const double2 *tw, double2 *work)
{
    hipLaunchKernelGGL(k_g_point, dim3(g1(M)), dim3(256), 0, c.stream, 0, work, U, chirp, Bhat, N, M);
    fft_grid(c, work, M, logM, tw, 0);
    hipLaunchKernelGGL(k_g_point, dim3(g1(M)), dim3(256), 0, c.stream, 1, work, U, chirp, Bhat, N, M);
    fft_grid(c, work, M, logM, tw, 1);
    hipLaunchKernelGGL(k_g_point, dim3(g1(M)), dim3(256), 0, c.stream, 2, work, U, chirp, Bhat, N, M);
}

static int smooth_giant(Ctx &c, int loop, int n)
{
    int Nmax = 2 * n - 2, logMmax = 0;
    while ((1LL << logMmax) < 2LL * Nmax - 1) ++logMmax;
    long long Mmax = 1LL << logMmax;
    size_t need = sizeof(double2) * (size_t)(3LL * n + Mmax / 2 + 2LL * n + 2 * Mmax + 2LL * n + 8);
    OVG_TRY(c.sm_giant.ensure(need));
    OVG_TRY(c.sm_gstate.ensure(sizeof(GiantState)));
    double2 *base = c.sm_giant.as<double2>();
    double *xs0 = (double *)base;
    double2 *tw = base + 3 * (size_t)n;
    double2 *chirp = tw + Mmax / 2;
    double2 *Bhat = chirp + 2 * (size_t)n;
    double2 *work = Bhat + Mmax;
    double2 *U = work + Mmax;
    GiantState *gs = c.sm_gstate.as<GiantState>();
    hipLaunchKernelGGL(k_giant_unwrap, dim3(1), dim3(kFftBlock), 0, c.stream, loop,
                       c.loop_off.as<int>(), c.loop_slots.as<int>(), c.x[0].as<double>(),
                       c.x[1].as<double>(), c.x[2].as<double>(), dparams(c), xs0, gs);
    GiantState h;
    OVG_HIP(hipMemcpyAsync(&h, gs, sizeof h, hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    const int N = h.span ? 2 * n - 2 : n;
    const int s = (int)(c.params.freq_to_cutoff * N / 2);
    if (N > 2 * s + 1 && N >= 2) {
        const int half = N / 2;
        int kr = half - s;
        int ki = (N % 2 == 0) ? half - s : half + 2 - s;
        if (ki > half + 1) ki = half + 1;
        if (N % 2 == 0 && ki > half) ki = half;
        int logM = 0;
        while ((1LL << logM) < 2LL * N - 1) ++logM;
        int M = 1 << logM;
        hipLaunchKernelGGL(k_g_tables, dim3(1024), dim3(256), 0, c.stream, N, M, tw, chirp, Bhat);
        fft_grid(c, Bhat, M, logM, tw, 0);
        for (int cc = 0; cc < 3; ++cc) {
            double *xc = xs0 + 2 * (size_t)n * cc;
            hipLaunchKernelGGL(k_g_coord, dim3(g1(N)), dim3(256), 0, c.stream, 0, xc, U, n, N, kr, ki);
            bluestein_grid(c, U, N, M, logM, chirp, Bhat, tw, work);
            hipLaunchKernelGGL(k_g_coord, dim3(g1(N)), dim3(256), 0, c.stream, 1, xc, U, n, N, kr, ki);
            bluestein_grid(c, U, N, M, logM, chirp, Bhat, tw, work);
            hipLaunchKernelGGL(k_g_coord, dim3(g1(n)), dim3(256), 0, c.stream, 2, xc, U, n, N, kr, ki);
        }
    }
    hipLaunchKernelGGL(k_giant_finish, dim3(g1(n)), dim3(256), 0, c.stream, loop,
                       c.loop_off.as<int>(), c.loop_slots.as<int>(), xs0, gs, c.x[0].as<double>(),
                       c.x[1].as<double>(), c.x[2].as<double>());
    return OVG_OK;
}

// Loop table from the host link mirror: loops in order of their smallest
// slot, each listed forward from it (src/vortex_smooth.c:192-242).
static int build_loops(Ctx &c)
{
    if (c.loops_version == c.topo_version) return OVG_OK;
    int n = c.n;
    std::vector<int> off, slots;
    std::vector<char> seen(n, 0);
    off.push_back(0);
    for (int i = 0; i < n; ++i) {
        if (seen[i] || c.h_status[2 * i] == OVG_EMPTY) continue;
        int k = i;
        do {
            seen[k] = 1;
            slots.push_back(k);
            k = c.h_conn[2 * k];
        } while (k != i && k >= 0 && !seen[k]);
        off.push_back((int)slots.size());
    }
    c.n_loops = (int)off.size() - 1;
    c.h_loop_off = off;
    OVG_TRY(c.loop_off.ensure(sizeof(int) * off.size()));
    OVG_TRY(c.loop_slots.ensure(sizeof(int) * (slots.size() + 1)));
    OVG_HIP(hipMemcpy(c.loop_off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    if (!slots.empty())
        OVG_HIP(hipMemcpy(c.loop_slots.p, slots.data(), sizeof(int) * slots.size(),
                          hipMemcpyHostToDevice));
    c.loops_version = c.topo_version;
    return OVG_OK;
}

int smooth_loops(Ctx &c)
{
    OVG_TRY(build_loops(c));
    if (c.n_loops == 0) return OVG_OK;
    size_t need = sizeof(double) * ((size_t)15 * c.n_live + (size_t)6 * c.n_loops + 64);
    OVG_TRY(c.sm_buf.ensure(need));
    std::vector<BigLoop> big;
    std::vector<int> giant;
    long long foff = 0;
    for (int lp = 0; lp < c.n_loops; ++lp) {
        int n = c.h_loop_off[lp + 1] - c.h_loop_off[lp];
        if (n <= kSmallLoop) continue;
        if (n > kMediumLoop) {
            giant.push_back(lp);
            continue;
        }
        int N = 2 * n - 2, logM = 0;
        while ((1 << logM) < 2 * N - 1) ++logM;
        int M = 1 << logM;
        big.push_back(BigLoop{lp, logM, foff});
        foff += 3LL * n + M / 2 + 2LL * n + 2LL * M + 2LL * n + 8;
    }
    hipLaunchKernelGGL(k_smooth, dim3(c.n_loops), dim3(kSmBlock), 0, c.stream, c.n_loops,
                       c.loop_off.as<int>(), c.loop_slots.as<int>(), c.x[0].as<double>(),
                       c.x[1].as<double>(), c.x[2].as<double>(), dparams(c),
                       c.params.freq_to_cutoff, c.sm_buf.as<double>());
    if (!big.empty()) {
        OVG_TRY(c.sm_fft.ensure(sizeof(double2) * (size_t)foff));
        OVG_TRY(c.sm_big.ensure(sizeof(BigLoop) * big.size()));
        OVG_HIP(hipMemcpyAsync(c.sm_big.p, big.data(), sizeof(BigLoop) * big.size(),
                               hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_smooth_fft, dim3((unsigned)big.size()), dim3(kFftBlock), 0,
                           c.stream, (int)big.size(), c.sm_big.as<BigLoop>(),
                           c.loop_off.as<int>(), c.loop_slots.as<int>(), c.x[0].as<double>(),
                           c.x[1].as<double>(), c.x[2].as<double>(), dparams(c),
                           c.params.freq_to_cutoff, c.sm_fft.as<double2>());
    }
    for (int lp : giant) OVG_TRY(smooth_giant(c, lp, c.h_loop_off[lp + 1] - c.h_loop_off[lp]));
    OVG_HIP(hipStreamSynchronize(c.stream));  // `big` is a host vector
    return OVG_OK;
}

}  // namespace ovg
