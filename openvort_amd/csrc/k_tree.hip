// Barnes-Hut octree build (src/octree.c:26-286) on the GPU.
//
// 1. k_keys: each live point descends from the root box (= domain box,
//    src/octree.c:150) through 21 levels using the reference's own child-box
//    arithmetic (make_child_box, src/octree.c:165-232) and its first-match
//    rule over the closed child boxes in the order BLF BLB BRF BRB TLF TLB
//    TRF TRB (src/octree.c:116-127).  The child indices form a 63-bit path
//    key, most significant level first, so key order is the reference's
//    traversal order.  Points outside the root box get bit 63 set (kept in
//    the root's sums, absent from every child: src/octree.c:116-128), with a
//    Morton code of their wrapped position below it for target coherence.
//    The tuned traversals (group, list) clamp points into the box and take
//    the same digits from integer Morton codes (morton_key).
// 2. A stable LSD radix sort of (key, slot), 8-bit digits, written here:
//    per-tile LDS histograms, one exclusive scan, then a scatter that ranks
//    equal digits with 64-lane ballots.
// 3. Compressed octree in DFS pre-order: the reference splits every cell
//    with more than one point (src/octree.c:112) and only descends into
//    nonempty children, so a chain of cells holding the same point set
//    collapses into its deepest cell without changing any acceptance
//    decision (Lm shrinks along the chain, com and circulation do not
//    change).  For sorted keys with adjacent common-prefix levels delta[i],
//    the nodes that start at point i are the distinct running minima of
//    delta[i..] above delta[i-1], plus i's leaf; a node's DFS index is
//    (nodes starting before i) + (rank among nodes starting at i) and its
//    skip rope is the first node starting after its range.
// 4. Sums (centre of mass, circulation).  Reference tree: bottom-up over
//    nodes bucketed by cell level, a node adding its children's sums in the
//    reference's order.  Tuned trees: a cell is a contiguous range of sorted
//    points, so its sums are differences of exact fixed-point prefix sums,
//    taken as the node is written (and again on a refit).  A leaf takes its
//    own point in both.
// HBM/latency-bound; no MFMA.
#include <algorithm>
#include <cmath>

#include "engine_internal.h"

namespace ovg {

constexpr uint64_t kDropKey = ~0ULL;

// make_child_box, src/octree.c:165-232, exact arithmetic per child index.
__device__ __forceinline__ void child_box(const double lo[3], const double hi[3], int idx,
                                          double nlo[3], double nhi[3])
{
    double c[3], h[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = lo[k] + hi[k];
        c[k] = c[k] * 0.5;
        h[k] = (hi[k] - lo[k]) / 2;
    }
    switch (idx) {
    case 0:
        nlo[0] = lo[0] + h[0]; nlo[1] = lo[1]; nlo[2] = lo[2];
        nhi[0] = c[0] + h[0]; nhi[1] = c[1]; nhi[2] = c[2];
        break;
    case 1:
        nlo[0] = lo[0]; nlo[1] = lo[1]; nlo[2] = lo[2];
        nhi[0] = c[0]; nhi[1] = c[1]; nhi[2] = c[2];
        break;
    case 2:
        nlo[0] = c[0]; nlo[1] = c[1]; nlo[2] = c[2] - h[2];
        nhi[0] = hi[0]; nhi[1] = hi[1]; nhi[2] = hi[2] - h[2];
        break;
    case 3:
        nlo[0] = lo[0]; nlo[1] = lo[1] + h[1]; nlo[2] = lo[2];
        nhi[0] = c[0]; nhi[1] = c[1] + h[1]; nhi[2] = c[2];
        break;
    case 4:
        nlo[0] = c[0]; nlo[1] = c[1] - h[1]; nlo[2] = c[2];
        nhi[0] = hi[0]; nhi[1] = hi[1] - h[1]; nhi[2] = hi[2];
        break;
    case 5:
        nlo[0] = lo[0]; nlo[1] = lo[1]; nlo[2] = lo[2] + h[2];
        nhi[0] = c[0]; nhi[1] = c[1]; nhi[2] = c[2] + h[2];
        break;
    case 6:
        nlo[0] = c[0]; nlo[1] = c[1]; nlo[2] = c[2];
        nhi[0] = hi[0]; nhi[1] = hi[1]; nhi[2] = hi[2];
        break;
    default:
        nlo[0] = c[0] - h[0]; nlo[1] = c[1]; nlo[2] = c[2];
        nhi[0] = hi[0] - h[0]; nhi[1] = hi[1]; nhi[2] = hi[2];
        break;
    }
}

// child_box without a branch on the child index: per axis each case takes
// one of four forms, (lo, c), (lo+h, c+h), (c, hi) or (c-h, hi-h); the form
// codes of the eight cases (2 bits per axis, 6 per case) sit in kChildForms.
// The values are the same expressions as in child_box.
constexpr uint64_t kChildForms =
    (1ull << 0) | (0ull << 2) | (0ull << 4) |                 // 0: B A A
    ((0ull | 0ull << 2 | 0ull << 4) << 6) |                   // 1: A A A
    ((2ull | 2ull << 2 | 3ull << 4) << 12) |                  // 2: C C D
    ((0ull | 1ull << 2 | 0ull << 4) << 18) |                  // 3: A B A
    ((2ull | 3ull << 2 | 2ull << 4) << 24) |                  // 4: C D C
    ((0ull | 0ull << 2 | 1ull << 4) << 30) |                  // 5: A A B
    ((2ull | 2ull << 2 | 2ull << 4) << 36) |                  // 6: C C C
    ((3ull | 2ull << 2 | 2ull << 4) << 42);                   // 7: D C C

__device__ __forceinline__ void child_box_sel(double lo[3], double hi[3], int idx)
{
    for (int k = 0; k < 3; ++k) {
        double c = lo[k] + hi[k];
        c = c * 0.5;
        const double h = (hi[k] - lo[k]) / 2;
        const int f = (int)((kChildForms >> (6 * idx + 2 * k)) & 3);
        const double nl = f == 0 ? lo[k] : f == 1 ? lo[k] + h : f == 2 ? c : c - h;
        const double nh = f == 0 ? c : f == 1 ? c + h : f == 2 ? hi[k] : hi[k] - h;
        lo[k] = nl;
        hi[k] = nh;
    }
}

__device__ __forceinline__ bool in_closed(const double lo[3], const double hi[3], double x,
                                          double y, double z)
{
    return x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2];
}

// Tuned modes: the cell of every level from the position quantised to 2^21
// steps of the root box, as 3-bit digits in the reference's child numbering
// (x upper half -> 0, y upper -> 2, z upper -> 4); these are the same dyadic
// cells, except that a point on a split plane goes to the upper cell instead
// of the reference's first match.
__device__ __forceinline__ uint64_t spread3(uint64_t v)  // bit b -> bit 3b
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Tuned keys resolve kTunedLevels levels (cells of 2^-16 of the box, 3e-6
// cm on a 0.2 cm box): the digits of deeper levels are zero, so the radix
// sort skips their bits, and points sharing a deepest cell sit in it as
// single-point leaves (that cell's Lm is its true size, so acceptance is
// never coarser than with deeper cells).
constexpr int kTunedLevels = 16;
// The 48 Morton bits sit at key bits 1..48 (level l's digit at 49 - 3l ..
// 51 - 3l), the tangle index above them (key_tangle).
constexpr int kTunedLowBit = 63 - 3 * kTunedLevels;  // where morton_key puts them
constexpr int kTunedDown = kTunedLowBit - 1;         // shift down to bit 1

// tangle << 49 | Morton bits of the clamped position in the tangle's box
__device__ __forceinline__ uint64_t morton_key(const TangleParams &P, double x, double y, double z)
{
    const double p3[3] = {x, y, z};
    constexpr double cells = (double)(1 << kTunedLevels);
    uint64_t q[3];
    for (int a = 0; a < 3; ++a) {
        double u = (p3[a] - P.lo[a]) / P.L[a] * cells;
        u = fmin(fmax(u, 0.0), cells - 1.0);
        q[a] = (uint64_t)u;
    }
    q[0] = ~q[0] & ((1u << kTunedLevels) - 1);
    // level l's digit comes from coordinate bit b = 21 - l (of a 21-bit
    // coordinate, here q << 5) and sits at key bits 63 - 3l = 3b (x),
    // 3b + 1 (y), 3b + 2 (z)
    constexpr int sh = kMaxLevel - kTunedLevels;
    const uint64_t m = spread3(q[0] << sh) | spread3(q[1] << sh) << 1 | spread3(q[2] << sh) << 2;
    return m >> kTunedDown;
}

// clamp != 0 (tuned modes): a point outside the root box is placed by its
// position clamped onto the box, so it stays in the tree and in Morton order
// among its neighbours; the reference's drop (src/octree.c:116-128) is kept
// for the reference-tree mode only.
__global__ void k_keys(int n_live, const int *__restrict__ live, const double *__restrict__ X,
                       const double *__restrict__ Y, const double *__restrict__ Z,
                       const TangleParams *__restrict__ Pp, uint64_t *__restrict__ keys,
                       int *__restrict__ vals, int *__restrict__ counters, int clamp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int k = live[i];
    const TangleParams &P = tparams_of(Pp, k);
    double x = X[k], y = Y[k], z = Z[k];
    if (clamp) {
        x = fmin(fmax(x, P.lo[0]), P.hi[0]);
        y = fmin(fmax(y, P.lo[1]), P.hi[1]);
        z = fmin(fmax(z, P.lo[2]), P.hi[2]);
    }
    double lo[3] = {P.lo[0], P.lo[1], P.lo[2]}, hi[3] = {P.hi[0], P.hi[1], P.hi[2]};
    uint64_t key = kDropKey;
    if (clamp) {
        const uint64_t t = Pp[0].n_tangles > 1 ? (uint64_t)Pp[0].tid[k] : 0;
        key = t << kKeyTangleShift | morton_key(P, x, y, z);
    } else if (!in_closed(lo, hi, x, y, z)) {
        // outside the root box: in no child (src/octree.c:116-128).  Bit 63
        // puts it after every in-box point; the remaining bits are a Morton
        // code of its wrapped position so that the traversal waves of these
        // targets stay spatially coherent.
        double p3[3] = {x, y, z};
        uint64_t m = 0;
        unsigned q[3];
        for (int a = 0; a < 3; ++a) {
            double u = (p3[a] - P.lo[a]) / P.L[a];
            u = u - floor(u);
            if (!(u >= 0.0 && u < 1.0)) u = 0.0;  // non-finite positions
            q[a] = (unsigned)(u * 1048576.0);
            if (q[a] > 1048575u) q[a] = 1048575u;
        }
        for (int b = 19; b >= 0; --b)
            for (int a = 0; a < 3; ++a) m = (m << 1) | ((q[a] >> b) & 1u);
        key = (1ULL << 63) | m;
    } else {
        key = 0;
        int fallbacks = 0;
        for (int lev = 1; lev <= kMaxLevel; ++lev) {
            // fast path: the point is far (relative to the box) from every
            // split plane, so exactly one child can hold it
            double c0 = (lo[0] + hi[0]) * 0.5, c1 = (lo[1] + hi[1]) * 0.5,
                   c2 = (lo[2] + hi[2]) * 0.5;
            double m0 = (hi[0] - lo[0]) * 1e-6, m1 = (hi[1] - lo[1]) * 1e-6,
                   m2 = (hi[2] - lo[2]) * 1e-6;
            int digit = -1;
            if (fabs(x - c0) > m0 && fabs(y - c1) > m1 && fabs(z - c2) > m2 &&
                fabs(x - lo[0]) > m0 && fabs(x - hi[0]) > m0 && fabs(y - lo[1]) > m1 &&
                fabs(y - hi[1]) > m1 && fabs(z - lo[2]) > m2 && fabs(z - hi[2]) > m2) {
                int xh = x > c0, yh = y > c1, zh = z > c2;
                digit = (xh ? 0 : 1) | (yh << 1) | (zh << 2);
            } else {
                double nlo[3], nhi[3];
                for (int ch = 0; ch < 8; ++ch) {
                    child_box(lo, hi, ch, nlo, nhi);
                    if (in_closed(nlo, nhi, x, y, z)) { digit = ch; break; }
                }
                if (digit < 0) {  // within an ulp of a face both boxes miss
                    int xh = x >= c0, yh = y > c1, zh = z > c2;
                    digit = (xh ? 0 : 1) | (yh << 1) | (zh << 2);
                    ++fallbacks;
                }
            }
            child_box_sel(lo, hi, digit);
            key |= (uint64_t)digit << (63 - 3 * lev);
        }
        if (fallbacks) atomicAdd(&counters[2], fallbacks);
    }
    keys[i] = key;
    vals[i] = k;
}

// counters[0] = in-box points: out-of-box keys have bit 63 set and sort
// last, so this is the first sorted index with the bit (n_live if none).
// Same-address atomics across the chip would serialise; this pass has none.
__global__ void k_count_inbox(int n_live, const uint64_t *__restrict__ keys,
                              int *__restrict__ counters)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n_live) return;
    const bool out = i < n_live && (keys[i] >> 63);
    const bool prev_in = i == 0 || !(keys[i - 1] >> 63);
    if ((out || i == n_live) && prev_in) counters[0] = i;
}

// ---------------------------------------------------------------- scan ----
constexpr int kScanThreads = 256, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

__device__ int block_exclusive_scan(int v, int *sh, int &total)
{
    // Hillis-Steele over 256 threads in LDS
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int a = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    int incl = sh[t];
    total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

// in and out may alias (in-place scan): every thread reads its items before
// the block scan's barriers and writes them after.
__global__ void k_scan_tile(const int *in, int *out, int n, int *__restrict__ tile_sums)
{
    __shared__ int sh[kScanThreads];
    int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = base + j < n ? in[base + j] : 0;
        s += v[j];
    }
    int total;
    int ex = block_exclusive_scan(s, sh, total);
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Each block adds the sum of the tile sums before it (a block reduction of
// at most kScanDirect values) to its tile; the last block also writes the
// total.  With more tiles the tile sums are first scanned in place (the
// same scan, one level up) and each block adds its own entry.
constexpr int kScanDirect = 2048;

__global__ void k_scan_add(int *out, int n, const int *__restrict__ sums, int tiles,
                           int *__restrict__ d_total, int scanned)
{
    __shared__ int red[kScanThreads / 64];
    int add = 0;
    if (scanned) {
        add = sums[blockIdx.x];
    } else {
        int pre = 0;
        for (int j = threadIdx.x; j < (int)blockIdx.x; j += kScanThreads) pre += sums[j];
        for (int off = 32; off > 0; off >>= 1) pre += __shfl_xor(pre, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pre;
        __syncthreads();
        for (int w = 0; w < kScanThreads / 64; ++w) add += red[w];
    }
    const int base = blockIdx.x * kScanTile;
    for (int j = threadIdx.x; j < kScanTile; j += kScanThreads)
        if (base + j < n) out[base + j] += add;
    if (!scanned && d_total && (int)blockIdx.x == tiles - 1 && threadIdx.x == 0)
        *d_total = add + sums[blockIdx.x];
}

// tmp holds the tile sums of every level (sized by scan_tmp_ints)
static size_t scan_tmp_ints(int n)
{
    size_t need = 0;
    for (int t = (n + kScanTile - 1) / kScanTile; t > 0; t = t > kScanDirect ? (t + kScanTile - 1) / kScanTile : 0)
        need += (size_t)t;
    return need + 1;
}

static void scan_level(Ctx &c, const int *d_in, int *d_out, int n, int *d_total, int *tmp)
{
    const int tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tile, dim3(tiles), dim3(kScanThreads), 0, c.stream, d_in, d_out,
                       n, tmp);
    if (tiles <= kScanDirect) {
        hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanThreads), 0, c.stream, d_out, n,
                           tmp, tiles, d_total, 0);
    } else {
        scan_level(c, tmp, tmp, tiles, d_total, tmp + tiles);
        hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanThreads), 0, c.stream, d_out, n,
                           tmp, tiles, nullptr, 1);
    }
}

__global__ void k_zero_ints(int *a, int n);

int tuned_key_cell_shift(int level)
{
    // two tuned keys share at least `level` levels (lcp_levels) exactly when
    // they agree above this bit
    return 64 - (63 - 3 * kTunedLevels) - 3 * level;
}

int launch_exclusive_scan(Ctx &c, const int *d_in, int *d_out, int n, int *d_total,
                          DBuf &tmp)
{
    if (n <= 0) {
        // a kernel, not a memset: callers may be captured into a HIP graph
        if (d_total) hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c.stream, d_total, 1);
        return OVG_OK;
    }
    OVG_TRY(tmp.ensure(sizeof(int) * scan_tmp_ints(n)));
    scan_level(c, d_in, d_out, n, d_total, tmp.as<int>());
    return OVG_OK;
}

// ---------------------------------------------------------- radix sort ----
constexpr int kRsThreads = 256, kRsRounds = 8, kRsTile = kRsThreads * kRsRounds;

__global__ void k_rs_hist(const uint64_t *__restrict__ keys, int n, int shift,
                          int *__restrict__ hist, int nblocks)
{
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsRounds; ++r) {
        int i = base + r * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// One tile of kRsTile keys per block, in input order (round, wave, lane):
// ranks among equal digits from 64-lane ballots, per-(round, wave) digit
// counts in LDS scanned per digit, then the tile is laid out by digit in LDS
// and written with consecutive threads on consecutive output positions of
// each digit's run.  Stable.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint64_t *__restrict__ kin, const int *__restrict__ vin, uint64_t *__restrict__ kout,
    int *__restrict__ vout, int n, int shift, const int *__restrict__ offs, int nblocks)
{
    constexpr int kSlots = kRsRounds * (kRsThreads / 64);
    __shared__ unsigned short cnt[kSlots][256];  // counts, then exclusive prefix per digit
    __shared__ int dstart[256];                  // tile-local start of each digit
    __shared__ int gbase[256];                   // this block's output offset per digit
    __shared__ int wsum[kRsThreads / 64];
    __shared__ uint64_t skey[kRsTile];
    __shared__ int sval[kRsTile];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int j = t; j < kSlots * 256; j += kRsThreads) (&cnt[0][0])[j] = 0;
    gbase[t] = offs[t * nblocks + blockIdx.x];
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const int base = blockIdx.x * kRsTile;
    uint64_t key[kRsRounds];
    int val[kRsRounds], dig[kRsRounds], rk[kRsRounds];
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        const int i = base + r * kRsThreads + t;
        const bool valid = i < n;
        key[r] = valid ? kin[i] : 0;
        val[r] = valid ? vin[i] : 0;
        dig[r] = valid ? (int)((key[r] >> shift) & 255) : 0;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((dig[r] >> b) & 1);
            peers &= ((dig[r] >> b) & 1) ? m : ~m;
        }
        rk[r] = __popcll(peers & below);
        const int leader = __ffsll((unsigned long long)peers) - 1;
        if (valid && lane == leader) cnt[r * (kRsThreads / 64) + w][dig[r]] = __popcll(peers);
    }
    __syncthreads();
    // thread t = digit t: exclusive prefix over the slots, then over digits
    int tot = 0;
    for (int q = 0; q < kSlots; ++q) {
        const int c = cnt[q][t];
        cnt[q][t] = (unsigned short)tot;
        tot += c;
    }
    int incl = tot;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int pre = 0;
    for (int q = 0; q < w; ++q) pre += wsum[q];
    dstart[t] = pre + incl - tot;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        const int i = base + r * kRsThreads + t;
        if (i < n) {
            const int d = dig[r];
            const int lp = dstart[d] + cnt[r * (kRsThreads / 64) + w][d] + rk[r];
            skey[lp] = key[r];
            sval[lp] = val[r];
        }
    }
    __syncthreads();
    const int nvalid = min(kRsTile, n - base);
    for (int j = t; j < nvalid; j += kRsThreads) {
        const uint64_t k = skey[j];
        const int d = (int)((k >> shift) & 255);
        const int pos = gbase[d] + j - dstart[d];
        kout[pos] = k;
        vout[pos] = sval[j];
    }
}

// Sorts c.tree.keys[0]/vals[0] (n entries) in place (result back in [0]).
int radix_sort(Ctx &c, int n, int first_shift, int last_shift)
{
    TreeBufs &T = c.tree;
    int nblocks = (n + kRsTile - 1) / kRsTile;
    OVG_TRY(T.hist.ensure(sizeof(int) * 256 * (size_t)nblocks));
    int src = 0;
    for (int shift = first_shift; shift <= last_shift; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(kRsThreads), 0, c.stream,
                           T.keys[src].as<uint64_t>(), n, shift, T.hist.as<int>(), nblocks);
        OVG_TRY(launch_exclusive_scan(c, T.hist.as<int>(), T.hist.as<int>(), 256 * nblocks,
                                      nullptr, T.scan_tmp));
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(kRsThreads), 0, c.stream,
                           T.keys[src].as<uint64_t>(), T.vals[src].as<int>(),
                           T.keys[1 - src].as<uint64_t>(), T.vals[1 - src].as<int>(), n, shift,
                           T.hist.as<int>(), nblocks);
        src = 1 - src;
    }
    if (src != 0) {
        OVG_HIP(hipMemcpyAsync(T.keys[0].p, T.keys[1].p, sizeof(uint64_t) * n,
                               hipMemcpyDeviceToDevice, c.stream));
        OVG_HIP(hipMemcpyAsync(T.vals[0].p, T.vals[1].p, sizeof(int) * n,
                               hipMemcpyDeviceToDevice, c.stream));
    }
    return OVG_OK;
}

// ----------------------------------------------------------- topology ----
// Common levels of two keys; equal keys share every level the keys resolve
// (max_lv: kMaxLevel for the reference tree, kTunedLevels for tuned trees).
// Tuned keys of two tangles share no level, not even the root: -1.
__device__ __forceinline__ int lcp_levels(uint64_t a, uint64_t b, int max_lv)
{
    if (a == b) return max_lv;
    if (max_lv == kTunedLevels) {
        if ((a ^ b) >> kKeyTangleShift) return -1;
        return (__clzll((long long)(a ^ b)) - (63 - 3 * kTunedLevels)) / 3;
    }
    return (__clzll((long long)(a ^ b)) - 1) / 3;
}

// sorted gather: slot ids, positions and pwrapped segment vectors in sorted
// order; delta[i] = common levels of sorted keys i and i+1 (in-box part).
__global__ void k_sorted_gather(int n_live, const uint64_t *__restrict__ keys,
                                const int *__restrict__ vals, const double *__restrict__ X,
                                const double *__restrict__ Y, const double *__restrict__ Z,
                                const int *__restrict__ fwd, const TangleParams *__restrict__ Pp,
                                const int *__restrict__ counters, int *__restrict__ spts,
                                double *__restrict__ sx, double *__restrict__ sy,
                                double *__restrict__ sz, double *__restrict__ qx,
                                double *__restrict__ qy, double *__restrict__ qz,
                                int8_t *__restrict__ delta, int *__restrict__ rank, int max_lv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int n_in = counters[0];
    int k = vals[i];
    const TangleParams &P = tparams_of(Pp, k);
    spts[i] = k;
    rank[k] = i;
    double x = X[k], y = Y[k], z = Z[k];
    sx[i] = x; sy[i] = y; sz[i] = z;
    int f = fwd[k];
    double ax = 0, ay = 0, az = 0;
    if (f >= 0) {  // octree_update_means circulation term, src/octree.c:276-281
        double bx = X[f], by = Y[f], bz = Z[f];
        pwrap(P, x, y, z, bx, by, bz);
        ax = bx - x; ay = by - y; az = bz - z;
    }
    qx[i] = ax; qy[i] = ay; qz[i] = az;
    if (i + 1 < n_in) delta[i] = (int8_t)lcp_levels(keys[i], keys[i + 1], max_lv);
}

// last j in [i+1, n_in-1] with lcp(K[i], K[j]) >= v (at least i+1
// qualifies).  The qualifying j form a run from i + 1, so a galloping search
// (steps 1, 2, 4, .. then bisection of the last step) takes about
// 2 log2(node size) key loads: most nodes hold a few points.
__device__ __forceinline__ int range_end(const uint64_t *__restrict__ keys, int i, int n_in,
                                         int v, int max_lv)
{
    const uint64_t ki = keys[i];
    int lo = i + 1, hi = n_in - 1;
    for (int step = 1; lo < hi; step <<= 1) {
        const int j = lo + step < hi ? lo + step : hi;
        if (lcp_levels(ki, keys[j], max_lv) >= v) {
            lo = j;
        } else {
            hi = j - 1;
            break;
        }
    }
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (lcp_levels(ki, keys[mid], max_lv) >= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------- tuned trees: prefix sums ----
// A cell of the tuned tree holds a contiguous range [a, b) of sorted points,
// so its sums are differences of prefix sums over the sorted points.  They
// are taken in 64-bit fixed point (units of 2^-shift cm, positions relative
// to the box corner), which makes every difference exact and independent of
// summation order.  Component 6 counts points whose values are not finite or
// lie beyond the fixed-point range (more than 4 box lengths out); a cell
// holding one gets NaN sums, as a floating-point sum would.
// In a forest every tangle has its own units (its own size and point
// count), so prefix sums over several tangles can pass 2^63: they are kept
// modulo 2^64 (unsigned), and a cell's sums, differences within one tangle,
// are exact.
constexpr int kFix = 7;
struct Fix7 {
    unsigned long long v[kFix];
};
constexpr int kFixThreads = 256, kFixItems = 4, kFixTile = kFixThreads * kFixItems;
constexpr int kFixScanThreads = 1024;

// Largest shift with n values of magnitude up to 4 Lmax summing below 2^62.
static int fix_shift(int n, double Lmax)
{
    const double lim = std::ldexp(1.0, 62) / ((double)(n > 0 ? n : 1) * 4.0 * Lmax);
    int k = (int)std::floor(std::log2(lim));
    return std::max(-60, std::min(60, k));
}

__device__ __forceinline__ unsigned long long to_fix(double v, double scale, double lim,
                                                     unsigned long long &bad)
{
    const double s = v * scale;
    if (!(fabs(s) <= lim)) {
        bad = 1;
        return 0;
    }
    return (unsigned long long)__double2ll_rn(s);
}

// Wave-inclusive scan of kFix 64-bit values (modulo 2^64).
__device__ __forceinline__ void wave_scan_fix(unsigned long long s[kFix], int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
        for (int f = 0; f < kFix; ++f) {
            const unsigned long long u = __shfl_up(s[f], off, 64);
            if (lane >= off) s[f] += u;
        }
}

// P[i] (i <= n) = fixed-point sums of sorted points [tile start, i); T[t] =
// the tile's total.  Tiles of kFixTile entries.  A point's units are its
// tangle's (tangle from its key in a forest, else the arguments).
__global__ __launch_bounds__(kFixThreads) void k_fix_tile(
    int n, const double *__restrict__ sx, const double *__restrict__ sy,
    const double *__restrict__ sz, const double *__restrict__ qx, const double *__restrict__ qy,
    const double *__restrict__ qz, const uint64_t *__restrict__ keys,
    const TangleParams *__restrict__ Pp, double lo0, double lo1, double lo2, double scale,
    double lim, Fix7 *__restrict__ P, Fix7 *__restrict__ T)
{
    __shared__ unsigned long long wsum[kFixThreads / 64][kFix];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int base = blockIdx.x * kFixTile + threadIdx.x * kFixItems;
    const bool forest = Pp[0].n_tangles > 1;
    unsigned long long v[kFixItems][kFix], s[kFix];
    for (int f = 0; f < kFix; ++f) s[f] = 0;
#pragma unroll
    for (int j = 0; j < kFixItems; ++j) {
        const int i = base + j;
        unsigned long long bad = 0;
        if (i < n) {
            double l0 = lo0, l1 = lo1, l2 = lo2, sc = scale, li = lim;
            if (forest) {
                const TangleParams &Q = Pp[key_tangle(keys[i])];
                l0 = Q.lo[0]; l1 = Q.lo[1]; l2 = Q.lo[2];
                sc = Q.fix_scale; li = Q.fix_lim;
            }
            v[j][0] = to_fix(sx[i] - l0, sc, li, bad);
            v[j][1] = to_fix(sy[i] - l1, sc, li, bad);
            v[j][2] = to_fix(sz[i] - l2, sc, li, bad);
            v[j][3] = to_fix(qx[i], sc, li, bad);
            v[j][4] = to_fix(qy[i], sc, li, bad);
            v[j][5] = to_fix(qz[i], sc, li, bad);
        } else {
            for (int f = 0; f < 6; ++f) v[j][f] = 0;
        }
        v[j][6] = bad;
        for (int f = 0; f < kFix; ++f) s[f] += v[j][f];
    }
    unsigned long long incl[kFix];
    for (int f = 0; f < kFix; ++f) incl[f] = s[f];
    wave_scan_fix(incl, lane);
    if (lane == 63)
        for (int f = 0; f < kFix; ++f) wsum[w][f] = incl[f];
    __syncthreads();
    unsigned long long ex[kFix], tot[kFix];
    for (int f = 0; f < kFix; ++f) {
        unsigned long long pre = 0, all = 0;
        for (int q = 0; q < kFixThreads / 64; ++q) {
            if (q < w) pre += wsum[q][f];
            all += wsum[q][f];
        }
        ex[f] = pre + incl[f] - s[f];
        tot[f] = all;
    }
#pragma unroll
    for (int j = 0; j < kFixItems; ++j) {
        const int i = base + j;
        if (i <= n)
            for (int f = 0; f < kFix; ++f) P[i].v[f] = ex[f];
        for (int f = 0; f < kFix; ++f) ex[f] += v[j][f];
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < kFix; ++f) T[blockIdx.x].v[f] = tot[f];
}

// Exclusive scan of the tile totals in place, one block: each thread takes a
// run of consecutive tiles.
__global__ __launch_bounds__(kFixScanThreads) void k_fix_tiles(int tiles, Fix7 *__restrict__ T)
{
    __shared__ unsigned long long wsum[kFixScanThreads / 64][kFix];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int per = (tiles + kFixScanThreads - 1) / kFixScanThreads;
    const int t0 = threadIdx.x * per, t1 = min(t0 + per, tiles);
    unsigned long long s[kFix];
    for (int f = 0; f < kFix; ++f) s[f] = 0;
    for (int t = t0; t < t1; ++t)
        for (int f = 0; f < kFix; ++f) s[f] += T[t].v[f];
    unsigned long long incl[kFix];
    for (int f = 0; f < kFix; ++f) incl[f] = s[f];
    wave_scan_fix(incl, lane);
    if (lane == 63)
        for (int f = 0; f < kFix; ++f) wsum[w][f] = incl[f];
    __syncthreads();
    unsigned long long ex[kFix];
    for (int f = 0; f < kFix; ++f) {
        unsigned long long pre = 0;
        for (int q = 0; q < w; ++q) pre += wsum[q][f];
        ex[f] = pre + incl[f] - s[f];
    }
    for (int t = t0; t < t1; ++t)
        for (int f = 0; f < kFix; ++f) {
            const unsigned long long u = T[t].v[f];
            T[t].v[f] = ex[f];
            ex[f] += u;
        }
}

// Sums of sorted points [a, b) (one tangle) into a node: centre of mass and
// circulation, from the prefix sums in units of 1/inv_scale cm.
__device__ __forceinline__ void fix_node_sums(Node &nd, const Fix7 *__restrict__ P,
                                              const Fix7 *__restrict__ T, int a, int b,
                                              const double lo[3], double inv_scale)
{
    const Fix7 &pa = P[a], &pb = P[b];
    const Fix7 &ta = T[a / kFixTile], &tb = T[b / kFixTile];
    long long s[kFix];
    for (int f = 0; f < kFix; ++f)
        s[f] = (long long)((pb.v[f] + tb.v[f]) - (pa.v[f] + ta.v[f]));
    const double nan = __builtin_nan("");
    const double inv = 1.0 / (double)(b - a);
    const bool bad = s[6] != 0;
    nd.cx = bad ? nan : lo[0] + (double)s[0] * inv_scale * inv;
    nd.cy = bad ? nan : lo[1] + (double)s[1] * inv_scale * inv;
    nd.cz = bad ? nan : lo[2] + (double)s[2] * inv_scale * inv;
    nd.qx = bad ? nan : (double)s[3] * inv_scale;
    nd.qy = bad ? nan : (double)s[4] * inv_scale;
    nd.qz = bad ? nan : (double)s[5] * inv_scale;
}

// Fixed-point units of a tangle of n points in a box of largest side Lmax:
// 2^shift per cm with n values of up to 4 Lmax summing below 2^62.
void fix_units(int n, double Lmax, double &scale, double &inv, double &lim)
{
    const int shift = fix_shift(n, Lmax);
    scale = std::ldexp(1.0, shift);
    inv = std::ldexp(1.0, -shift);
    lim = std::ldexp(1.0, 62) / (double)std::max(n, 1);
}

// Fixed-point prefix sums of the sorted points (c.tree.sx .. sqz).  One
// tangle: units from its live points and box (inv_scale returned); a
// forest: each tangle's units from its parameter entry.
static int fix_prefix(Ctx &c, int n, double &inv_scale)
{
    TreeBufs &T = c.tree;
    const ovg_params &p = c.params;
    double Lmax = 0;
    for (int k = 0; k < 3; ++k) Lmax = std::max(Lmax, p.box_hi[k] - p.box_lo[k]);
    double scale, lim;
    fix_units(n, Lmax, scale, inv_scale, lim);
    const int tiles = (n + 1 + kFixTile - 1) / kFixTile;
    OVG_TRY(T.fix.ensure(sizeof(Fix7) * ((size_t)n + 1)));
    OVG_TRY(T.fix_tiles.ensure(sizeof(Fix7) * (size_t)tiles));
    hipLaunchKernelGGL(k_fix_tile, dim3(tiles), dim3(kFixThreads), 0, c.stream, n,
                       T.sx.as<double>(), T.sy.as<double>(), T.sz.as<double>(),
                       T.sqx.as<double>(), T.sqy.as<double>(), T.sqz.as<double>(),
                       T.keys[0].as<uint64_t>(), dparams(c), p.box_lo[0], p.box_lo[1],
                       p.box_lo[2], scale, lim, T.fix.as<Fix7>(), T.fix_tiles.as<Fix7>());
    hipLaunchKernelGGL(k_fix_tiles, dim3(1), dim3(kFixScanThreads), 0, c.stream, tiles,
                       T.fix_tiles.as<Fix7>());
    return OVG_OK;
}

// Walk the nodes starting at sorted point i; EMIT=false only counts them.
// fixP != nullptr (tuned trees): a cell's Lm is the root's largest side
// times 2^-level and its sums come from the prefix sums; otherwise Lm
// replays the reference's child boxes and the sums run level by level
// afterwards.  Leaves take their point either way.  Tuned trees of a
// forest: every tangle's first sorted point starts that tangle's root (its
// box, its points), whose DFS index goes to roots[tangle].
template <bool EMIT>
__global__ void k_nodes(int n_live, const uint64_t *__restrict__ keys,
                        const int8_t *__restrict__ delta, const int *__restrict__ counters,
                        int *__restrict__ counts, const int *__restrict__ base,
                        Node *__restrict__ nodes, NodeMeta *__restrict__ meta,
                        const TangleParams *__restrict__ Pp, const double *__restrict__ sx,
                        const double *__restrict__ sy, const double *__restrict__ sz,
                        const double *__restrict__ qx, const double *__restrict__ qy,
                        const double *__restrict__ qz, const Fix7 *__restrict__ fixP,
                        const Fix7 *__restrict__ fixT, double inv_scale, int max_lv,
                        int *__restrict__ roots)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int n_in = counters[0];
    if (i >= n_in) {  // dropped points start no node
        if (!EMIT) counts[i] = i == 0 ? 1 : 0;
        if (EMIT && i == 0) {
            // no in-box point at all: the root is the only node
            const TangleParams &P = *Pp;
            Node r{};
            r.count = n_live;
            r.skip = base[n_live];
            double Lx = P.hi[0] - P.lo[0], Ly = P.hi[1] - P.lo[1], Lz = P.hi[2] - P.lo[2];
            r.Lm = Lx > Ly ? (Lx > Lz ? Lx : Lz) : (Ly > Lz ? Ly : Lz);
            nodes[0] = r;
            NodeMeta m{};
            m.start = 0; m.level = 0;
            meta[0] = m;
            if (roots) roots[0] = 0;
        }
        return;
    }
    int dl = i == 0 ? -1 : delta[i - 1];
    int dr = i == n_in - 1 ? -1 : delta[i];
    int cnt = 0;
    int d = EMIT ? base[i] : 0;
    int total = EMIT ? base[n_live] : 0;
    uint64_t ki = keys[i];
    const bool forest = fixP && Pp[0].n_tangles > 1;
    const int tg = forest ? key_tangle(ki) : 0;
    const TangleParams &P = Pp[tg];
    const double fixinv = forest ? P.fix_inv : inv_scale;
    const double Lroot = fmax(fmax(P.hi[0] - P.lo[0], P.hi[1] - P.lo[1]), P.hi[2] - P.lo[2]);
    const double plo[3] = {P.lo[0], P.lo[1], P.lo[2]};
    auto emit = [&](int level, int end, int count) {
        if (EMIT) {
            Node nd{};
            nd.count = count;
            int nxt = end + 1;
            nd.skip = nxt < n_in ? base[nxt] : total;
            if (count >= 2 && level >= 1) {
                int lv = level < kMaxLevel ? level : kMaxLevel;
                if (fixP) {
                    nd.Lm = ldexp(Lroot, -lv);
                    fix_node_sums(nd, fixP, fixT, i, i + count, plo, fixinv);
                } else {
                    // replay the reference's child boxes down to `level`
                    double lo[3] = {P.lo[0], P.lo[1], P.lo[2]};
                    double hi[3] = {P.hi[0], P.hi[1], P.hi[2]};
                    for (int l = 1; l <= lv; ++l)
                        child_box_sel(lo, hi, (int)((ki >> (63 - 3 * l)) & 7));
                    double Lx = hi[0] - lo[0], Ly = hi[1] - lo[1], Lz = hi[2] - lo[2];
                    nd.Lm = Lx > Ly ? (Lx > Lz ? Lx : Lz) : (Ly > Lz ? Ly : Lz);
                }
            }
            if (count == 1) {  // a leaf's sums are its point (octree_update_means)
                nd.cx = sx[i]; nd.cy = sy[i]; nd.cz = sz[i];
                nd.qx = qx[i]; nd.qy = qy[i]; nd.qz = qz[i];
            }
            nodes[d] = nd;
            NodeMeta m{};
            m.start = i;
            m.level = level;
            meta[d] = m;
            ++d;
        }
        ++cnt;
    };
    if (dl < 0) {
        // explicit root: the domain box with every live point of the tangle
        // (src/octree.c:141-150); one tangle: every live point, the dropped
        // ones included
        const int end = !forest ? n_live - 1 : dr < 0 ? i : range_end(keys, i, n_in, 0, max_lv);
        if (EMIT) {
            Node r{};
            r.count = end - i + 1;
            r.skip = end + 1 < n_in ? base[end + 1] : total;
            double Lx = P.hi[0] - P.lo[0], Ly = P.hi[1] - P.lo[1], Lz = P.hi[2] - P.lo[2];
            r.Lm = Lx > Ly ? (Lx > Lz ? Lx : Lz) : (Ly > Lz ? Ly : Lz);
            if (fixP) fix_node_sums(r, fixP, fixT, i, end + 1, plo, fixinv);
            if (r.count == 1) {
                r.cx = sx[i]; r.cy = sy[i]; r.cz = sz[i];
                r.qx = qx[i]; r.qy = qy[i]; r.qz = qz[i];
            }
            nodes[d] = r;
            NodeMeta m{};
            m.start = i; m.level = 0;
            meta[d] = m;
            if (roots) roots[tg] = d;
            ++d;
        }
        ++cnt;
    }
    int v = dl + 1;
    if (v < 1) v = 1;
    while (v <= dr) {
        int e = range_end(keys, i, n_in, v, max_lv);
        int w = lcp_levels(ki, keys[e], max_lv);
        emit(w, e, e - i + 1);
        v = w + 1;
    }
    int leaf = (dl > dr ? dl : dr) + 1;
    if (leaf < 1) leaf = 1;
    emit(leaf, i, 1);
    if (!EMIT) counts[i] = cnt;
}

__device__ void sum_node(int d, const int *__restrict__ counters, Node *__restrict__ nodes,
                         const NodeMeta *__restrict__ meta, const double *__restrict__ sx,
                         const double *__restrict__ sy, const double *__restrict__ sz,
                         const double *__restrict__ qx, const double *__restrict__ qy,
                         const double *__restrict__ qz, int n_live)
{
    Node nd = nodes[d];
    int n_in = counters[0];
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (d == 0) {
        for (int c = 1; c < nd.skip; c = nodes[c].skip) {
            s[0] += nodes[c].cx; s[1] += nodes[c].cy; s[2] += nodes[c].cz;
            s[3] += nodes[c].qx; s[4] += nodes[c].qy; s[5] += nodes[c].qz;
        }
        for (int j = n_in; j < n_live; ++j) {
            s[0] += sx[j]; s[1] += sy[j]; s[2] += sz[j];
            s[3] += qx[j]; s[4] += qy[j]; s[5] += qz[j];
        }
    } else if (nd.count == 1) {
        int j = meta[d].start;
        s[0] = sx[j]; s[1] = sy[j]; s[2] = sz[j];
        s[3] = qx[j]; s[4] = qy[j]; s[5] = qz[j];
    } else {
        for (int c = d + 1; c < nd.skip; c = nodes[c].skip) {
            s[0] += nodes[c].cx; s[1] += nodes[c].cy; s[2] += nodes[c].cz;
            s[3] += nodes[c].qx; s[4] += nodes[c].qy; s[5] += nodes[c].qz;
        }
    }
    nd.cx = s[0]; nd.cy = s[1]; nd.cz = s[2];
    nd.qx = s[3]; nd.qy = s[4]; nd.qz = s[5];
    nodes[d] = nd;
}


// Internal nodes (and the root) bucketed by cell level, levels 0 ..
// kMaxLevel + 1, without atomics: per-block counts in hist[level][block], an
// exclusive scan over that array, then a scatter at block offset + rank in
// the block.  The order inside a level is free (a level's nodes only read
// deeper ones); leaves already hold their sums.
constexpr int kLevels = kMaxLevel + 2;

__device__ __forceinline__ int sum_level(const Node *__restrict__ nodes,
                                         const NodeMeta *__restrict__ meta, int d)
{
    return (d == 0 || nodes[d].count > 1) ? meta[d].level : -1;
}

__global__ void k_level_count(const int *__restrict__ counters, const Node *__restrict__ nodes,
                              const NodeMeta *__restrict__ meta, int *__restrict__ hist)
{
    __shared__ int cnt[kLevels];
    if (threadIdx.x < kLevels) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = d < counters[1] ? sum_level(nodes, meta, d) : -1;
    if (l >= 0) atomicAdd(&cnt[l], 1);
    __syncthreads();
    if (threadIdx.x < kLevels) hist[threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void k_level_scatter(const int *__restrict__ counters, const Node *__restrict__ nodes,
                                const NodeMeta *__restrict__ meta,
                                const int *__restrict__ offs, int *__restrict__ list)
{
    __shared__ int cnt[kLevels];
    if (threadIdx.x < kLevels) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = d < counters[1] ? sum_level(nodes, meta, d) : -1;
    if (l < 0) return;
    list[offs[l * gridDim.x + blockIdx.x] + atomicAdd(&cnt[l], 1)] = d;
}

// Sums bottom-up for the nodes of one cell level: a leaf takes its point,
// an internal node adds its children (d+1, then along skip ropes); the
// root also adds the points outside the root box.  Sums stay in cx..qz
// until k_finish_means turns cx..cz into centres of mass.  A fixed grid
// strides over the level's bucket.
__global__ void k_sums_level(int level, int nblk, const int *__restrict__ counters,
                             const int *__restrict__ offs, const int *__restrict__ list,
                             Node *__restrict__ nodes, const NodeMeta *__restrict__ meta,
                             const double *__restrict__ sx, const double *__restrict__ sy,
                             const double *__restrict__ sz, const double *__restrict__ qx,
                             const double *__restrict__ qy, const double *__restrict__ qz,
                             int n_live)
{
    const int b0 = offs[level * nblk],
              b1 = level + 1 < kLevels ? offs[(level + 1) * nblk] : offs[kLevels * nblk];
    for (int e = b0 + blockIdx.x * blockDim.x + threadIdx.x; e < b1;
         e += gridDim.x * blockDim.x)
        sum_node(list[e], counters, nodes, meta, sx, sy, sz, qx, qy, qz, n_live);
}
// com = sum * (1.0/N), src/octree.c:269; leaves keep x exactly.
__global__ void k_finish_means(const int *__restrict__ counters, Node *__restrict__ nodes)
{
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= counters[1]) return;
    Node &nd = nodes[d];
    if (nd.count > 1) {
        double inv = 1.0 / nd.count;
        nd.cx = nd.cx * inv;
        nd.cy = nd.cy * inv;
        nd.cz = nd.cz * inv;
    }
}

__global__ void k_zero_ints(int *a, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0;
}

__global__ void k_tree_counters(int n_live, const int *__restrict__ base, int *__restrict__ counters)
{
    counters[1] = base[n_live];
}

int build_tree(Ctx &c, const double *const x[3])
{
    TreeBufs &T = c.tree;
    int n = c.n_live;
    ++c.tree_id;
    size_t cap_nodes = 2 * (size_t)n + 2;
    for (int b = 0; b < 2; ++b) {
        OVG_TRY(T.keys[b].ensure(sizeof(uint64_t) * (n + 1)));
        OVG_TRY(T.vals[b].ensure(sizeof(int) * (n + 1)));
    }
    OVG_TRY(T.delta.ensure(n + 1));
    OVG_TRY(T.counts.ensure(sizeof(int) * (n + 1)));
    OVG_TRY(T.base.ensure(sizeof(int) * (n + 2)));
    OVG_TRY(T.nodes.ensure(sizeof(Node) * cap_nodes));
    OVG_TRY(T.meta.ensure(sizeof(NodeMeta) * cap_nodes));
    OVG_TRY(T.level_hist.ensure(sizeof(int) * ((size_t)kLevels * ((cap_nodes + 255) / 256) + 1)));
    OVG_TRY(T.level_list.ensure(sizeof(int) * cap_nodes));
    OVG_TRY(T.counters.ensure(sizeof(int) * 64));
    OVG_TRY(T.roots.ensure(sizeof(int) * (size_t)std::max(c.n_tangles, 1)));
    OVG_TRY(T.sorted_pts.ensure(sizeof(int) * (n + 1)));
    OVG_TRY(T.rank.ensure(sizeof(int) * ((size_t)c.n + 1)));
    DBuf *sv[6] = {&T.sx, &T.sy, &T.sz, &T.sqx, &T.sqy, &T.sqz};
    for (auto *b : sv) OVG_TRY(b->ensure(sizeof(double) * (n + 1)));
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c.stream, T.counters.as<int>(), 64);
    if (n == 0) return launch_status();
    int bs = 256, gs = (n + bs - 1) / bs;
    int clamp = c.params.traverse != OVG_TRAVERSE_POINT;
    hipLaunchKernelGGL(k_keys, dim3(gs), dim3(bs), 0, c.stream, n, c.live.as<int>(), x[0], x[1],
                       x[2], dparams(c), T.keys[0].as<uint64_t>(), T.vals[0].as<int>(),
                       T.counters.as<int>(), clamp);
    // tuned keys: Morton bits 1 .. 48 in 8-bit digits (6 passes), and the
    // tangle index above them in a forest (1 or 2 more)
    const int max_lv = clamp ? kTunedLevels : kMaxLevel;
    const int last = c.n_tangles > 256 ? 57 : c.n_tangles > 1 ? 49 : 41;
    if (clamp) OVG_TRY(radix_sort(c, n, 1, last));
    else OVG_TRY(radix_sort(c, n, 0, 56));
    hipLaunchKernelGGL(k_count_inbox, dim3((n + 1 + bs - 1) / bs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), T.counters.as<int>());
    hipLaunchKernelGGL(k_sorted_gather, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), T.vals[0].as<int>(), x[0], x[1], x[2],
                       c.fwd.as<int>(), dparams(c), T.counters.as<int>(), T.sorted_pts.as<int>(),
                       T.sx.as<double>(), T.sy.as<double>(), T.sz.as<double>(),
                       T.sqx.as<double>(), T.sqy.as<double>(), T.sqz.as<double>(),
                       T.delta.as<int8_t>(), T.rank.as<int>(), max_lv);
    // tuned trees (every point in the box): sums from prefix sums as the
    // nodes are written
    double inv_scale = 0;
    if (clamp) OVG_TRY(fix_prefix(c, n, inv_scale));
    const Fix7 *fixP = clamp ? T.fix.as<Fix7>() : nullptr;
    const Fix7 *fixT = clamp ? T.fix_tiles.as<Fix7>() : nullptr;
    hipLaunchKernelGGL(k_nodes<false>, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), T.delta.as<int8_t>(), T.counters.as<int>(),
                       T.counts.as<int>(), T.base.as<int>(), T.nodes.as<Node>(),
                       T.meta.as<NodeMeta>(), dparams(c), T.sx.as<double>(), T.sy.as<double>(),
                       T.sz.as<double>(), T.sqx.as<double>(), T.sqy.as<double>(),
                       T.sqz.as<double>(), fixP, fixT, inv_scale, max_lv, T.roots.as<int>());
    // base[i] for i <= n; counts[n] is not written -> scan n entries, total at base[n]
    OVG_TRY(launch_exclusive_scan(c, T.counts.as<int>(), T.base.as<int>(), n,
                                  T.base.as<int>() + n, T.scan_tmp));
    hipLaunchKernelGGL(k_nodes<true>, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), T.delta.as<int8_t>(), T.counters.as<int>(),
                       T.counts.as<int>(), T.base.as<int>(), T.nodes.as<Node>(),
                       T.meta.as<NodeMeta>(), dparams(c), T.sx.as<double>(), T.sy.as<double>(),
                       T.sz.as<double>(), T.sqx.as<double>(), T.sqy.as<double>(),
                       T.sqz.as<double>(), fixP, fixT, inv_scale, max_lv, T.roots.as<int>());
    hipLaunchKernelGGL(k_tree_counters, dim3(1), dim3(1), 0, c.stream, n, T.base.as<int>(),
                       T.counters.as<int>());
    if (clamp) return OVG_OK;
    // reference tree: sums level by level in the reference's order
    int gn = (int)((cap_nodes + bs - 1) / bs);
    int *hist = T.level_hist.as<int>(), *list = T.level_list.as<int>();
    hipLaunchKernelGGL(k_level_count, dim3(gn), dim3(bs), 0, c.stream, T.counters.as<int>(),
                       T.nodes.as<Node>(), T.meta.as<NodeMeta>(), hist);
    OVG_TRY(launch_exclusive_scan(c, hist, hist, kLevels * gn, hist + kLevels * gn, T.scan_tmp));
    hipLaunchKernelGGL(k_level_scatter, dim3(gn), dim3(bs), 0, c.stream, T.counters.as<int>(),
                       T.nodes.as<Node>(), T.meta.as<NodeMeta>(), hist, list);
    const int gl = std::min(gn, 1024);
    for (int level = kMaxLevel + 1; level >= 0; --level)
        hipLaunchKernelGGL(k_sums_level, dim3(gl), dim3(bs), 0, c.stream, level, gn,
                           T.counters.as<int>(), hist, list, T.nodes.as<Node>(),
                           T.meta.as<NodeMeta>(), T.sx.as<double>(), T.sy.as<double>(),
                           T.sz.as<double>(), T.sqx.as<double>(), T.sqy.as<double>(),
                           T.sqz.as<double>(), n);
    hipLaunchKernelGGL(k_finish_means, dim3(gn), dim3(bs), 0, c.stream, T.counters.as<int>(),
                       T.nodes.as<Node>());
    return OVG_OK;
}

// Refit (tuned trees): every node's sums from the prefix sums at the new
// positions; leaves (count 1, not the root) take their point.
__global__ void k_refit_nodes(const int *__restrict__ counters, Node *__restrict__ nodes,
                              const NodeMeta *__restrict__ meta, const double *__restrict__ sx,
                              const double *__restrict__ sy, const double *__restrict__ sz,
                              const double *__restrict__ qx, const double *__restrict__ qy,
                              const double *__restrict__ qz, const Fix7 *__restrict__ fixP,
                              const Fix7 *__restrict__ fixT, const TangleParams *__restrict__ Pp,
                              double inv_scale)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= counters[1]) return;
    Node &nd = nodes[d];
    const int j = meta[d].start, cnt = nd.count;
    if (cnt == 1) {  // a leaf (or a one-point root) takes its point
        nd.cx = sx[j]; nd.cy = sy[j]; nd.cz = sz[j];
        nd.qx = qx[j]; nd.qy = qy[j]; nd.qz = qz[j];
    } else {
        const TangleParams &P = *Pp;
        const double lo[3] = {P.lo[0], P.lo[1], P.lo[2]};
        fix_node_sums(nd, fixP, fixT, j, j + cnt, lo, inv_scale);
    }
}

// Same tree (keys, order, nodes, cell sizes) at new positions x: positions,
// segment vectors and node sums recomputed.  Tuned trees only (every point
// in the box, so the root's range is all sorted points).
int refit_tree(Ctx &c, const double *const x[3])
{
    TreeBufs &T = c.tree;
    const int n = c.n_live;
    if (n == 0) return OVG_OK;
    if (c.params.traverse == OVG_TRAVERSE_POINT)
        return set_error(OVG_EINVAL, "refit_tree: the reference tree is rebuilt, not refitted");
    const int bs = 256, gs = (n + bs - 1) / bs;
    const size_t cap_nodes = 2 * (size_t)n + 2;
    const int gn = (int)((cap_nodes + bs - 1) / bs);
    hipLaunchKernelGGL(k_sorted_gather, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), T.vals[0].as<int>(), x[0], x[1], x[2],
                       c.fwd.as<int>(), dparams(c), T.counters.as<int>(), T.sorted_pts.as<int>(),
                       T.sx.as<double>(), T.sy.as<double>(), T.sz.as<double>(),
                       T.sqx.as<double>(), T.sqy.as<double>(), T.sqz.as<double>(),
                       T.delta.as<int8_t>(), T.rank.as<int>(), kTunedLevels);
    double inv_scale = 0;
    OVG_TRY(fix_prefix(c, n, inv_scale));
    hipLaunchKernelGGL(k_refit_nodes, dim3(gn), dim3(bs), 0, c.stream, T.counters.as<int>(),
                       T.nodes.as<Node>(), T.meta.as<NodeMeta>(), T.sx.as<double>(),
                       T.sy.as<double>(), T.sz.as<double>(), T.sqx.as<double>(),
                       T.sqy.as<double>(), T.sqz.as<double>(), T.fix.as<Fix7>(),
                       T.fix_tiles.as<Fix7>(), dparams(c), inv_scale);
    return OVG_OK;
}

}  // namespace ovg
