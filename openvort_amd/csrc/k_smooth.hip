// Frequency-filter smoothing (smooth -> high_freq_cutoff -> low_pass_filter,
// src/vortex_smooth.c:134-175,183-453,459-462), one workgroup per closed
// loop.
//
// Per loop (in slot order of its first slot, src/vortex_smooth.c:192-210):
// thread 0 unwraps the coordinates along the loop with seg_pwrap against the
// previous unwrapped point (:229-242).  A loop whose ends are more than half
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
constexpr int kSmallLoop = 1024;

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
    if (n > kSmallLoop) return;  // FFT path
    const int *slots = loop_slots + off;
    // work layout per loop (15n + 6 doubles from 15*off + 6*lp): three
    // coordinate arrays of 2n, three coefficient arrays of 2n + 2, three
    // output arrays of n
    double *w = work + (size_t)15 * off + (size_t)6 * lp;
    double *xs[3] = {w, w + 2 * (size_t)n, w + 4 * (size_t)n};
    double *cf = w + 6 * (size_t)n;
    double *yo = cf + 3 * (2 * (size_t)n + 2);
    const size_t cstride = 2 * (size_t)n + 2;
    __shared__ int s_span;
    __shared__ double s_coef[3], s_x0[3];
    if (threadIdx.x == 0) {
        int k = slots[0];
        double px = X[k], py = Y[k], pz = Z[k];
        xs[0][0] = px; xs[1][0] = py; xs[2][0] = pz;
        for (int j = 1; j < n; ++j) {
            int m = slots[j];
            double qx = X[m], qy = Y[m], qz = Z[m];
            pwrap(P, px, py, pz, qx, qy, qz);
            xs[0][j] = qx; xs[1][j] = qy; xs[2][j] = qz;
            px = qx; py = qy; pz = qz;
        }
        double ex = xs[0][0] - xs[0][n - 1], ey = xs[1][0] - xs[1][n - 1],
               ez = xs[2][0] - xs[2][n - 1];
        double len = sqrt(dot3(ex, ey, ez, ex, ey, ez));
        s_span = len > (P.hi[0] - P.lo[0]) / 2;
    }
    __syncthreads();
    const bool span = s_span;
    if (span) {
        if (threadIdx.x < 3) {
            double *a = xs[threadIdx.x];
            s_coef[threadIdx.x] = (a[n - 1] - a[0]) / (n - 1);
            s_x0[threadIdx.x] = a[0];
        }
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += kSmBlock)
            for (int c = 0; c < 3; ++c) {
                double *a = xs[c];
                if (j == 0 || j == n - 1) continue;
                a[j] = a[j] - s_coef[c] * j - s_x0[c];
            }
        __syncthreads();
        if (threadIdx.x < 3) {
            xs[threadIdx.x][0] = 0;
            xs[threadIdx.x][n - 1] = 0;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < n - 2; j += kSmBlock)
            for (int c = 0; c < 3; ++c) xs[c][n + j] = -xs[c][n - 2 - j];
        __syncthreads();
    }
    const int N = span ? 2 * n - 2 : n;
    const int s = (int)(fcut * N / 2);
    if (N > 2 * s + 1 && N >= 2) {
        // surviving halfcomplex entries: r_k for 0 <= k < kr, i_k for 1 <= k < ki
        const int half = N / 2;
        const int kr = half - s;
        int ki = (N % 2 == 0) ? half - s : half + 2 - s;
        if (ki > half + 1) ki = half + 1;
        if (N % 2 == 0 && ki > half) ki = half;
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
    if (span) {
        for (int j = threadIdx.x; j < n; j += kSmBlock)
            for (int c = 0; c < 3; ++c) xs[c][j] = xs[c][j] + s_coef[c] * j + s_x0[c];
        __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += kSmBlock) {
        int m = slots[j];
        X[m] = xs[0][j];
        Y[m] = xs[1][j];
        Z[m] = xs[2][j];
    }
}

// ---- large loops: Bluestein transform over radix-2 FFTs ------------------
// For a loop of N > kSmallLoop coordinates the DFT U_k = sum_j u_j
// e^{-2 pi i jk/N} is computed as c_k (a * b)_k with the chirp
// c_j = e^{-pi i j^2/N}, a_j = u_j c_j, b_m = conj(c_m) (Bluestein 1970),
// the cyclic convolution of length M = 2^ceil(log2(2N-1)) done with
// radix-2 FFTs in the loop's global scratch by one 1024-thread workgroup.
// The filter then zeroes the same halfcomplex entries as the small-loop
// path, and y = conj(DFT(conj(U)))/N.
constexpr int kFftBlock = 1024;

struct BigLoop {
    int loop;       // index into the loop table
    int logM;
    long long off;  // first double2 of this loop's scratch
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// in place, forward (sign -1) with twiddles tw[k] = e^{-2 pi i k/M}, k < M/2
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
        int stride = M >> st;  // twiddle index step
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

// U = DFT(u) of length N through the precomputed Bhat = FFT(b); work has M
// entries.  u and U may alias.
__device__ void bluestein(const double2 *u, double2 *U, int N, int M, int logM,
                          const double2 *chirp, const double2 *Bhat, const double2 *tw,
                          double2 *work)
{
    for (int j = threadIdx.x; j < M; j += blockDim.x)
        work[j] = j < N ? cmul(u[j], chirp[j]) : make_double2(0, 0);
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
    const int M = 1 << B.logM;
    // scratch: xs (3 x 2n doubles = 3n double2), tw (M/2), chirp (2n), Bhat (M),
    // work (M), U (2n)
    double2 *base = scratch + B.off;
    double *xs[3] = {(double *)base, (double *)base + 2 * (size_t)n, (double *)base + 4 * (size_t)n};
    double2 *tw = base + 3 * (size_t)n;
    double2 *chirp = tw + M / 2;
    double2 *Bhat = chirp + 2 * (size_t)n;
    double2 *work = Bhat + M;
    double2 *U = work + M;
    __shared__ int s_span;
    __shared__ double s_coef[3], s_x0[3];
    if (threadIdx.x == 0) {
        int k = slots[0];
        double px = X[k], py = Y[k], pz = Z[k];
        xs[0][0] = px; xs[1][0] = py; xs[2][0] = pz;
        for (int j = 1; j < n; ++j) {
            int m = slots[j];
            double qx = X[m], qy = Y[m], qz = Z[m];
            pwrap(P, px, py, pz, qx, qy, qz);
            xs[0][j] = qx; xs[1][j] = qy; xs[2][j] = qz;
            px = qx; py = qy; pz = qz;
        }
        double ex = xs[0][0] - xs[0][n - 1], ey = xs[1][0] - xs[1][n - 1],
               ez = xs[2][0] - xs[2][n - 1];
        s_span = sqrt(dot3(ex, ey, ez, ex, ey, ez)) > (P.hi[0] - P.lo[0]) / 2;
    }
    __syncthreads();
    const bool span = s_span;
    if (span) {
        if (threadIdx.x < 3) {
            double *a = xs[threadIdx.x];
            s_coef[threadIdx.x] = (a[n - 1] - a[0]) / (n - 1);
            s_x0[threadIdx.x] = a[0];
        }
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x)
            for (int c = 0; c < 3; ++c) {
                if (j == 0 || j == n - 1) continue;
                xs[c][j] = xs[c][j] - s_coef[c] * j - s_x0[c];
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
    const int N = span ? 2 * n - 2 : n;
    const int s = (int)(fcut * N / 2);
    if (N > 2 * s + 1 && N >= 2) {
        const int half = N / 2;
        const int kr = half - s;
        int ki = (N % 2 == 0) ? half - s : half + 2 - s;
        if (ki > half + 1) ki = half + 1;
        if (N % 2 == 0 && ki > half) ki = half;
        // twiddles, chirp (exact argument reduction) and Bhat = FFT(b)
        for (int k = threadIdx.x; k < M / 2; k += blockDim.x) {
            double sn, cs;
            sincospi(-2.0 * (double)k / M, &sn, &cs);
            tw[k] = make_double2(cs, sn);
        }
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            long long q = ((long long)j * j) % (2LL * N);
            double sn, cs;
            sincospi(-(double)q / N, &sn, &cs);
            chirp[j] = make_double2(cs, sn);
        }
        __syncthreads();
        for (int m = threadIdx.x; m < M; m += blockDim.x) {
            double2 v = make_double2(0, 0);
            if (m < N) v = conj2(chirp[m]);
            else if (m > M - N) v = conj2(chirp[M - m]);
            Bhat[m] = v;
        }
        __syncthreads();
        fft_radix2(Bhat, M, B.logM, tw, false);
        for (int c = 0; c < 3; ++c) {
            for (int j = threadIdx.x; j < N; j += blockDim.x) U[j] = make_double2(xs[c][j], 0.0);
            __syncthreads();
            bluestein(U, U, N, M, B.logM, chirp, Bhat, tw, work);
            // keep the surviving halfcomplex entries, mirrored to k > N/2
            for (int k = threadIdx.x; k < N; k += blockDim.x) {
                int kk = k <= N - k ? k : N - k;
                double2 v = U[k];
                bool keep_r = kk < kr;
                bool keep_i = kk >= 1 && kk < ki;
                U[k] = conj2(make_double2(keep_r ? v.x : 0.0, keep_i ? v.y : 0.0));
            }
            __syncthreads();
            bluestein(U, U, N, M, B.logM, chirp, Bhat, tw, work);
            for (int j = threadIdx.x; j < n; j += blockDim.x) xs[c][j] = U[j].x / N;
            __syncthreads();
        }
    }
    if (span) {
        for (int j = threadIdx.x; j < n; j += blockDim.x)
            for (int c = 0; c < 3; ++c) xs[c][j] = xs[c][j] + s_coef[c] * j + s_x0[c];
        __syncthreads();
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        int m = slots[j];
        X[m] = xs[0][j];
        Y[m] = xs[1][j];
        Z[m] = xs[2][j];
    }
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
    // loops longer than kSmallLoop take the FFT path; their scratch follows
    std::vector<BigLoop> big;
    long long foff = 0;
    for (int lp = 0; lp < c.n_loops; ++lp) {
        int n = c.h_loop_off[lp + 1] - c.h_loop_off[lp];
        if (n <= kSmallLoop) continue;
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
        OVG_HIP(hipStreamSynchronize(c.stream));  // big is a host vector
    }
    return OVG_OK;
}

}  // namespace ovg
