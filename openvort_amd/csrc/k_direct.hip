// Direct Biot-Savart sum with periodic images (src/tangle.c:652-687).
//
// Each thread owns one target point; a block of 256 targets streams the
// source segments through LDS in tiles of 256 (r1, pwrapped r2, slot ids:
// 56 B each) and applies all 27 image shifts to the target, so every LDS
// read feeds 27 pair evaluations.  The source range is split over
// gridDim.y chunks for parallelism at small N; chunk partials are summed
// in chunk order by a second kernel, so results do not depend on timing.
// VALU-bound (no MFMA: the pair formula is not a contraction).
#include <algorithm>
#include <cstdint>
#include <mutex>

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
    const int stride_tgt = n_tgt;  // the partials' row length
    int i = blockIdx.x * kDirBlock + threadIdx.x;
    int k = -1;
    double rx = 0, ry = 0, rz = 0;
    if (i < n_tgt) {
        k = tgt[i];  // negative: a checked-out slot, skipped
        if (k >= 0) {
            rx = X[k]; ry = Y[k]; rz = Z[k];
        }
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
        size_t stride = (size_t)stride_tgt;
        double *o = part + (size_t)blockIdx.y * 6 * stride;
        o[0 * stride + i] = apx; o[1 * stride + i] = apy; o[2 * stride + i] = apz;
        o[3 * stride + i] = aix; o[4 * stride + i] = aiy; o[5 * stride + i] = aiz;
    }
}

// Sum chunk partials in chunk order and scatter to slot-indexed outputs
// (a negative slot is skipped).
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
    if (k < 0) return;
    px[k] = s[0]; py[k] = s[1]; pz[k] = s[2];
    ix[k] = s[3]; iy[k] = s[4]; iz[k] = s[5];
}

// Source chunks of one direct launch: enough blocks for >= 4 per CU on 256
// CUs, chunks of whole source tiles; at most max_chunks (>= 1).
static void direct_plan(const Ctx &c, int n_tgt, int max_chunks, int &chunks, int &chunk_len)
{
    const int tiles = (n_tgt + kDirBlock - 1) / kDirBlock;
    int want = (2048 + tiles - 1) / tiles;
    int maxc = (c.n_src + kDirBlock - 1) / kDirBlock;
    chunks = want < maxc ? want : maxc;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    chunk_len = (c.n_src + chunks - 1) / chunks;
    chunk_len = (chunk_len + kDirBlock - 1) / kDirBlock * kDirBlock;
    if (chunk_len == 0) chunk_len = kDirBlock;
    chunks = (c.n_src + chunk_len - 1) / chunk_len;
    if (chunks < 1) chunks = 1;
}

static void direct_launch(Ctx &c, const double *const x[3], int n_tgt, const int *d_tgt_slots,
                          double *const out_prim[3], double *const out_img[3], int chunks,
                          int chunk_len, double *part)
{
    const int tiles = (n_tgt + kDirBlock - 1) / kDirBlock;
    dim3 grid(tiles, chunks);
    bool mixed = c.params.precision == OVG_PREC_MIXED;
    if (mixed)
        hipLaunchKernelGGL(k_direct<true>, grid, dim3(kDirBlock), 0, c.stream, n_tgt,
                           d_tgt_slots, x[0], x[1], x[2], c.n_src, chunk_len,
                           c.seg[0].as<double>(), c.seg[1].as<double>(), c.seg[2].as<double>(),
                           c.seg[3].as<double>(), c.seg[4].as<double>(), c.seg[5].as<double>(),
                           c.seg_ids.as<int2>(), dparams(c), part);
    else
        hipLaunchKernelGGL(k_direct<false>, grid, dim3(kDirBlock), 0, c.stream, n_tgt,
                           d_tgt_slots, x[0], x[1], x[2], c.n_src, chunk_len,
                           c.seg[0].as<double>(), c.seg[1].as<double>(), c.seg[2].as<double>(),
                           c.seg[3].as<double>(), c.seg[4].as<double>(), c.seg[5].as<double>(),
                           c.seg_ids.as<int2>(), dparams(c), part);
    int bs = 256, gs = (n_tgt + bs - 1) / bs;
    hipLaunchKernelGGL(k_direct_reduce, dim3(gs), dim3(bs), 0, c.stream, n_tgt, chunks,
                       d_tgt_slots, part, out_prim[0], out_prim[1], out_prim[2], out_img[0],
                       out_img[1], out_img[2]);
}

int launch_direct(Ctx &c, const double *const x[3], int n_tgt, const int *d_tgt_slots,
                  double *const out_prim[3], double *const out_img[3], int chunks)
{
    if (n_tgt == 0) return OVG_OK;
    int chunk_len;
    if (chunks <= 0) {
        direct_plan(c, n_tgt, INT32_MAX, chunks, chunk_len);
    } else {
        chunk_len = (c.n_src + chunks - 1) / chunks;
        chunk_len = (chunk_len + kDirBlock - 1) / kDirBlock * kDirBlock;
        if (chunk_len == 0) chunk_len = kDirBlock;
        chunks = (c.n_src + chunk_len - 1) / chunk_len;
        if (chunks < 1) chunks = 1;
    }
    OVG_TRY(c.dpart.ensure(sizeof(double) * 6 * (size_t)n_tgt * chunks));
    direct_launch(c, x, n_tgt, d_tgt_slots, out_prim, out_img, chunks, chunk_len,
                  c.dpart.as<double>());
    return launch_status();
}

// ---- the allocation-free direct sum (ovg_velocity_at_noalloc) -------------
// Chunk partials and the checked target list live in one file-scope device
// array per GPU, `scratch`, held until the process ends.  Host code never
// uses the array's name as an address: its device address is resolved once
// per device (direct_scratch_init, at context creation) and kept.  Every
// call on a GPU shares its array, so calls that overlap on one GPU race.
// Layout: [0, 6 * m * chunks) doubles of partials, then m ints of slots.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
// direct_plan gives tiles * ceil(2048 / tiles) <= 2048 + tiles - 1 (tile,
// chunk) blocks, so at most this many (target, chunk) partials for
// OVG_NOALLOC_MAX_TARGETS targets: the scratch takes the plan that
// ovg_velocity_at uses (same chunks, same sums) for every allowed size
constexpr size_t kMaxTargetChunks =
    (size_t)kDirBlock * (2048 + OVG_NOALLOC_MAX_TARGETS / kDirBlock - 1);
static_assert(kScratchDoubles >= 6 * kMaxTargetChunks + OVG_NOALLOC_MAX_TARGETS / 2,
              "scratch holds the partials of the chunk rule's largest plan");
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxScratchDevices] = {};

int direct_scratch_init(int device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return set_error(OVG_ELIMIT, "device %d: scratch is kept for devices 0..%d", device,
                         kMaxScratchDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return OVG_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device)
        return set_error(OVG_EDEVICE, "scratch of device %d resolved with device %d current",
                         device, cur);
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(OVG_EDEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s",
                         device, hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
    return OVG_OK;
}

// Checked copy of the caller's target list: a slot out of range or EMPTY
// becomes -1 (the kernels skip it, its outputs are NaN) and sets bit 4 of
// the context's device flags.
__global__ void k_check_slots(int m, const int *__restrict__ in, int n,
                              const int *__restrict__ status, int *__restrict__ out,
                              int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int k = in[i];
    const bool ok = k >= 0 && k < n && status[k] != OVG_EMPTY;
    out[i] = ok ? k : -1;
    if (!ok) atomicOr(flags, 4);
}

int launch_direct_noalloc(Ctx &c, const double *const x[3], int n_tgt, const int *d_slots,
                          double *const out_prim[3], double *const out_img[3],
                          const int **checked)
{
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c.device >= 0 && c.device < kMaxScratchDevices) base = g_scratch_ptr[c.device];
    }
    if (!base)
        return set_error(OVG_EDEVICE, "scratch of device %d is not resolved", c.device);
    if (n_tgt > OVG_NOALLOC_MAX_TARGETS)
        return set_error(OVG_ELIMIT, "%d targets exceed OVG_NOALLOC_MAX_TARGETS (%d)", n_tgt,
                         OVG_NOALLOC_MAX_TARGETS);
    const size_t slot_doubles = ((size_t)n_tgt * sizeof(int) + 7) / 8;
    const int max_chunks = (int)((kScratchDoubles - slot_doubles) / (6 * (size_t)n_tgt));
    int chunks, chunk_len;
    direct_plan(c, n_tgt, max_chunks, chunks, chunk_len);
    int *slots = reinterpret_cast<int *>(base + 6 * (size_t)n_tgt * chunks);
    hipLaunchKernelGGL(k_check_slots, dim3((n_tgt + 255) / 256), dim3(256), 0, c.stream, n_tgt,
                       d_slots, c.n, c.status.as<int>(), slots, c.flags.as<int>());
    direct_launch(c, x, n_tgt, slots, out_prim, out_img, chunks, chunk_len, base);
    *checked = slots;
    return launch_status();
}

}  // namespace ovg
