// Direct Biot-Savart sum with periodic images (src/tangle.c:652-687).
//
// Each thread owns one target point; a block of 256 targets streams the
// source segments through LDS in tiles of 256 (r1, pwrapped r2, slot ids:
// 56 B each) and applies all 27 image shifts to the target, so every LDS
// read feeds 27 pair evaluations.  The source range is split over
// gridDim.y chunks for parallelism at small N; chunk partials are summed
// in chunk order by a second kernel, so results do not depend on timing.
// VALU-bound (no MFMA: the pair formula is not a contraction).
#include "engine_internal.h"

namespace ovg {

constexpr int kDirBlock = 256;

template <bool MIXED>
__global__ __launch_bounds__(kDirBlock) void k_direct(
    int n_tgt, const int *__restrict__ tgt, const double *__restrict__ X,
    const double *__restrict__ Y, const double *__restrict__ Z, int n_src, int chunk_len,
    const double *__restrict__ s1x, const double *__restrict__ s1y,
    const double *__restrict__ s1z, const double *__restrict__ s2x,
    const double *__restrict__ s2y, const double *__restrict__ s2z,
    const int2 *__restrict__ ids, const TangleParams *__restrict__ Pp,
    double *__restrict__ part)
{
    __shared__ double sh[6][kDirBlock];
    __shared__ int2 shid[kDirBlock];
    const TangleParams &P = *Pp;
    int i = blockIdx.x * kDirBlock + threadIdx.x;
    int k = -1;
    double rx = 0, ry = 0, rz = 0;
    if (i < n_tgt) {
        k = tgt[i];
        rx = X[k]; ry = Y[k]; rz = Z[k];
    }
    const int nimg = P.n_images;
    const double f = P.f_bs;
    double apx = 0, apy = 0, apz = 0, aix = 0, aiy = 0, aiz = 0;
    int s_begin = blockIdx.y * chunk_len;
    int s_end = min(n_src, s_begin + chunk_len);
    for (int t0 = s_begin; t0 < s_end; t0 += kDirBlock) {
        int j = t0 + threadIdx.x;
        __syncthreads();
        if (j < s_end) {
            sh[0][threadIdx.x] = s1x[j]; sh[1][threadIdx.x] = s1y[j];
            sh[2][threadIdx.x] = s1z[j]; sh[3][threadIdx.x] = s2x[j];
            sh[4][threadIdx.x] = s2y[j]; sh[5][threadIdx.x] = s2z[j];
            shid[threadIdx.x] = ids[j];
        }
        __syncthreads();
        int cnt = min(kDirBlock, s_end - t0);
        if (k < 0) continue;
        for (int s = 0; s < cnt; ++s) {
            double ax = sh[0][s], ay = sh[1][s], az = sh[2][s];
            double bx = sh[3][s], by = sh[4][s], bz = sh[5][s];
            int2 id = shid[s];
            // primary box: skip the two segments handled by LIA (:658)
            if (!(id.x == k || id.y == k)) {
                if (MIXED) {
                    seg_field_mixed(ax - rx, ay - ry, az - rz, bx - rx, by - ry, bz - rz, f,
                                    (float)f, apx, apy, apz);
                } else {
                    double vx, vy, vz;
                    if (seg_field64(ax - rx, ay - ry, az - rz, bx - rx, by - ry, bz - rz, f,
                                    vx, vy, vz)) {
                        apx += vx; apy += vy; apz += vz;
                    }
                }
            }
            // image boxes at shifted(r) (src/tangle.c:370-414); the direct
            // path adds them unreflected (src/tangle.c:678-683)
            for (int im = 0; im < nimg; ++im) {
                double qx, qy, qz;
                image_point(P, im, rx, ry, rz, qx, qy, qz);
                if (MIXED) {
                    seg_field_mixed(ax - qx, ay - qy, az - qz, bx - qx, by - qy, bz - qz, f,
                                    (float)f, aix, aiy, aiz);
                } else {
                    double vx, vy, vz;
                    if (seg_field64(ax - qx, ay - qy, az - qz, bx - qx, by - qy, bz - qz, f,
                                    vx, vy, vz)) {
                        aix += vx; aiy += vy; aiz += vz;
                    }
                }
            }
        }
    }
    if (i < n_tgt) {
        size_t stride = (size_t)n_tgt;
        double *o = part + (size_t)blockIdx.y * 6 * stride;
        o[0 * stride + i] = apx; o[1 * stride + i] = apy; o[2 * stride + i] = apz;
        o[3 * stride + i] = aix; o[4 * stride + i] = aiy; o[5 * stride + i] = aiz;
    }
}

// Sum chunk partials in chunk order and scatter to slot-indexed outputs.
__global__ void k_direct_reduce(int n_tgt, int chunks, const int *__restrict__ tgt,
                                const double *__restrict__ part, double *__restrict__ px,
                                double *__restrict__ py, double *__restrict__ pz,
                                double *__restrict__ ix, double *__restrict__ iy,
                                double *__restrict__ iz)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tgt) return;
    size_t stride = (size_t)n_tgt;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c) {
        const double *o = part + (size_t)c * 6 * stride;
        for (int q = 0; q < 6; ++q) s[q] += o[q * stride + i];
    }
    int k = tgt[i];
    px[k] = s[0]; py[k] = s[1]; pz[k] = s[2];
    ix[k] = s[3]; iy[k] = s[4]; iz[k] = s[5];
}

int launch_direct(Ctx &c, const double *const x[3], int n_tgt, const int *d_tgt_slots,
                  double *const out_prim[3], double *const out_img[3], int chunks)
{
    if (n_tgt == 0) return OVG_OK;
    int tiles = (n_tgt + kDirBlock - 1) / kDirBlock;
    if (chunks <= 0) {
        // aim for >= 4 blocks per CU on 256 CUs, chunks of >= 1 source tile
        int want = (2048 + tiles - 1) / tiles;
        int maxc = (c.n_src + kDirBlock - 1) / kDirBlock;
        chunks = want < maxc ? want : maxc;
        if (chunks < 1) chunks = 1;
    }
    int chunk_len = (c.n_src + chunks - 1) / chunks;
    chunk_len = (chunk_len + kDirBlock - 1) / kDirBlock * kDirBlock;
    if (chunk_len == 0) chunk_len = kDirBlock;
    chunks = (c.n_src + chunk_len - 1) / chunk_len;
    if (chunks < 1) chunks = 1;
    OVG_TRY(c.dpart.ensure(sizeof(double) * 6 * (size_t)n_tgt * chunks));
    dim3 grid(tiles, chunks);
    bool mixed = c.params.precision == OVG_PREC_MIXED;
    if (mixed)
        hipLaunchKernelGGL(k_direct<true>, grid, dim3(kDirBlock), 0, c.stream, n_tgt,
                           d_tgt_slots, x[0], x[1], x[2], c.n_src, chunk_len,
                           c.seg[0].as<double>(), c.seg[1].as<double>(), c.seg[2].as<double>(),
                           c.seg[3].as<double>(), c.seg[4].as<double>(), c.seg[5].as<double>(),
                           c.seg_ids.as<int2>(), dparams(c), c.dpart.as<double>());
    else
        hipLaunchKernelGGL(k_direct<false>, grid, dim3(kDirBlock), 0, c.stream, n_tgt,
                           d_tgt_slots, x[0], x[1], x[2], c.n_src, chunk_len,
                           c.seg[0].as<double>(), c.seg[1].as<double>(), c.seg[2].as<double>(),
                           c.seg[3].as<double>(), c.seg[4].as<double>(), c.seg[5].as<double>(),
                           c.seg_ids.as<int2>(), dparams(c), c.dpart.as<double>());
    int bs = 256, gs = (n_tgt + bs - 1) / bs;
    hipLaunchKernelGGL(k_direct_reduce, dim3(gs), dim3(bs), 0, c.stream, n_tgt, chunks,
                       d_tgt_slots, c.dpart.as<double>(), out_prim[0], out_prim[1],
                       out_prim[2], out_img[0], out_img[1], out_img[2]);
    return launch_status();
}

}  // namespace ovg
