// Barnes-Hut traversal with per-wavefront interaction lists
// (OVG_TRAVERSE_LIST).
//
// One wavefront (one 64-thread block) per group of PPW Morton-adjacent
// target points.  The wave walks the octree once for the whole group and
// all 27 copies (primary and images) together, each stack range carrying
// its copy index:
//   - a node is accepted as a monopole when its centre of mass lies farther
//     than Lm/theta from the group's bounding box (shifted by the image), so
//     every point of the group would accept it under the reference's
//     per-point test Lm/|r - com| < theta (src/octree.c:319);
//   - a single-point node is always taken as its segment (N == 1, :319);
//   - a node of at most leaf_size points that is not accepted contributes
//     each of its segments;
//   - anything else is opened.
// The walk is parallel over the wave's lanes: a stack in LDS holds ranges
// [first, end) of sibling nodes in DFS pre-order, tagged with their copy;
// each round every lane pops one range, handles its first node and pushes the sibling
// continuation [skip, end) and, when it opens the node, its children
// [d + 1, skip).  Accepted entries are compacted into LDS lists:
//   far entries (farther than mixed_near from the box, OVG_PREC_MIXED) as
//     FP32 offsets from the group centre plus the FP32 circulation;
//   near entries (and every entry in FP64 mode) as node / segment indices,
//     evaluated in FP64 with the reference's guards (src/tangle.c:549-552).
// A full list is evaluated by all 64 lanes at once: lane = (slice, target)
// with 64 / ny slices striding the list for a group of ny targets (groups
// end at cell boundaries, so many hold fewer than PPW), so nearly every lane
// does useful pair work and the walk is paid once per group.  The
// per-slice FP64 partial sums are added in slice order at the end, so the
// result does not depend on scheduling.  The primary and image parts are
// one sum (written as the primary part; the image part is zero).
//
// Cost model: VALU-bound in the pair evaluation (44 flops per pair);
// the walk costs about one node load per lane-round.
#include <algorithm>

#include "engine_internal.h"

namespace ovg {

constexpr int kListStack = 384;  // ranges per wave
constexpr int kListFar = 128;    // FP32 far entries per flush
constexpr int kListNear = 128;   // near entries per flush
// In the single-pop regime the stack grows by at most one pending sibling
// range per tree level below the popped node (22 levels: root + 21), plus
// the two pushes of the popped range itself.
constexpr int kStackMargin = 32;
constexpr int kRecCap = 6144;    // recorded entries per group (about 3000 at theta 0.4)
// near entry: bits 0-25 node or segment index, 26-30 copy (0 = primary,
// c = image c - 1), bit 31 set for a node
constexpr unsigned kIdxMask = (1u << 26) - 1;
constexpr int kCopyShift = 26;
constexpr unsigned kNodeBit = 1u << 31;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int prefix_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
}

// Inclusive wave prefix sum and wave total of small non-negative values
// (< 2^bits, bits wave-uniform), one ballot per bit: no cross-lane data
// movement through LDS.
__device__ __forceinline__ int wave_scan_small(int v, int bits, int &total)
{
    int below = 0, sum = 0;
    for (int b = 0; b < bits; ++b) {
        const unsigned long long m = __ballot((v >> b) & 1);
        below += prefix_below(m) << b;
        sum += __popcll(m) << b;
    }
    total = sum;
    return below + v;
}

__device__ __forceinline__ int wave_max_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ double wave_min_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_max_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// A wave-uniform double moved to scalar registers.
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Squared distance from p to the box [lo, hi].
__device__ __forceinline__ double box_dist2(double px, double py, double pz, double lx,
                                            double ly, double lz, double hx, double hy,
                                            double hz)
{
    double dx = fmax(fmax(lx - px, px - hx), 0.0);
    double dy = fmax(fmax(ly - py, py - hy), 0.0);
    double dz = fmax(fmax(lz - pz, pz - hz), 0.0);
    return dx * dx + dy * dy + dz * dz;
}

// Consecutive blocks are dispatched round-robin over the 8 XCDs; give each
// XCD a contiguous range of groups so neighbouring groups, which walk the
// same tree nodes, share one L2.
__device__ __forceinline__ int xcd_group(int b, int nb)
{
    const int x = b & 7, i = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// A mirrored copy's source (kMirror), one axis: segment (x, x + q)
// reflected on the mirrored axes (r), with sh = 2 * wall there and L * shift
// on the others; c is the reflected x.  With M the reflection, the
// reference's M v(M r + c) equals det(M) times the field of the reflected
// segments at r, so a copy mirrored on an odd number of axes also reverses
// its segments (rev): start x~ + q~, vector -q~.
__device__ __forceinline__ void mirror_axis(bool r, bool rev, double sh, double &x, double &q,
                                            double &c)
{
    const double qt = r ? -q : q;
    c = r ? sh - x : x + sh;
    x = rev ? c + qt : c;
    q = rev ? -qt : qt;
}

// MODE: kWalk, kRecord (walk and keep every group's accepted entries, in
// emission order, in rec[g * rec_cap ..], their count in rec_count[g], -1
// when they do not fit), kReplay (a group with a record evaluates those
// entries at the current positions instead of walking; the others walk).

template <bool MIXED, int PPW, int MODE>
__global__ __launch_bounds__(64, 4) void k_bh_list(
    unsigned *__restrict__ rec, int *__restrict__ rec_count, int rec_cap, int rec_groups,
    const int2 *__restrict__ groups, const int *__restrict__ spts, const double *__restrict__ X,
    const double *__restrict__ Y, const double *__restrict__ Z, const Node *__restrict__ nodes,
    const NodeMeta *__restrict__ meta, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ sqx, const double *__restrict__ sqy,
    const double *__restrict__ sqz, const int *__restrict__ counters,
    const int *__restrict__ n_groups_p, const TangleParams *__restrict__ Pp,
    const int *__restrict__ roots, double *__restrict__ px, double *__restrict__ py, double *__restrict__ pz,
    double *__restrict__ ix, double *__restrict__ iy, double *__restrict__ iz,
    unsigned long long *__restrict__ icount, int stack_cap, int far_tile)
{
    static_assert(PPW <= 64 && 4 * (64 / PPW) <= kListFar, "slices of a full group");
    // kMirror: copies may mirror.  A mirrored copy of segment (x, q) is the
    // segment reflected through the walls, reversed when the copy mirrors an
    // odd number of axes (the reference's reflected target and reflected
    // velocity, src/tangle.c:380-410,644-646, moved onto the source; see
    // mirror_axis); acceptance takes the reflected centre (2 w - x on a
    // mirrored axis, x + L s on the others).
    constexpr bool MIR = MODE == kMirror;
    static_assert(kListFar <= 128, "the far flush checks two entries per lane");
    __shared__ int2 stk[kListStack];
    __shared__ float4 farA[kListFar];  // offset a from the copy's group centre, qx
    __shared__ float4 farB[kListFar];  // qy, qz, (a x q)_x, (a x q)_y
    __shared__ float farC[kListFar];   // (a x q)_z
    __shared__ unsigned nearI[kListNear];
    __shared__ int bEnd[64], bStart[64], bCopy[64];  // bucket cells of one round
    __shared__ double shL[27][3];                    // copy shifts (2 * wall on mirrored axes)
    __shared__ int shM[27];                          // mirrored axes of each copy (kMirror)

    const int lane = lane_id();
    const int n_groups = *n_groups_p;
    // a forest's groups each lie in one tangle: its parameters, copy shifts
    // and root node (loaded again when the tangle changes)
    const bool forest = Pp[0].n_tangles > 1;
    int cur_tangle = -1;
    // a grid sized for more groups than the device counted (a capacity
    // bound) maps only its first n_groups blocks, so every XCD gets a share
    const int nb = min((int)gridDim.x, n_groups);
    if ((int)blockIdx.x >= nb) return;
    for (int g = xcd_group(blockIdx.x, nb); g < n_groups; g += nb) {
    const int2 grp = groups[g];
    const int tg = forest ? __builtin_amdgcn_readfirstlane(Pp[0].tid[spts[grp.x]]) : 0;
    const TangleParams &P = Pp[tg];
    if (tg != cur_tangle) {
        __syncthreads();
        if (lane < 27) {
            const bool img = lane >= 1 && lane <= P.n_images;
            const int m = MIR && img ? P.refl[lane - 1] : 0;
            // a zero made here: a constant hoisted out of the group loop
            // would stay live across it and be spilled
            double zero;
            asm volatile("v_mov_b64 %0, 0" : "=v"(zero));
            for (int q = 0; q < 3; ++q)
                shL[lane][q] = !img ? zero : (m >> q & 1) ? P.W2[lane - 1][q] : P.Ls[lane - 1][q];
            if (MIR) shM[lane] = m;
        }
        __syncthreads();
        cur_tangle = tg;
    }
    const double near2 = P.near2;
    // lane = (slice, target) over the group's ny targets: S = 64 / ny slices
    // (at most 32 and at most far_tile / 4, so a flush round of 4 S entries
    // fits the far list), the lanes from S * ny on idle.  A far flush holds
    // whole rounds of 4 S; far_tile (96 or 128 <= kListFar) leaves room for
    // a round of 64 entries and the remainder of a flush for every ny.
    const int ny = min(max(grp.y, 1), PPW);
    const int S = min(min(64 / ny, 32), far_tile / 4);
    const int far_cap = far_tile / (4 * S) * (4 * S);
    const bool valid = grp.y > 0 && lane < S * ny;
    const int slice = valid ? lane / ny : 0;
    const int t = grp.x + lane % ny;
    const int k = valid ? spts[t] : 0;
    const double rx = valid ? X[k] : 0.0, ry = valid ? Y[k] : 0.0, rz = valid ? Z[k] : 0.0;
    const bool on = valid && isfinite(rx) && isfinite(ry) && isfinite(rz);

    // group bounding box over the finite targets
    const double inf = __builtin_inf();
    const double lox = uniform(wave_min_d(on ? rx : inf)),
                 loy = uniform(wave_min_d(on ? ry : inf)),
                 loz = uniform(wave_min_d(on ? rz : inf));
    const double hix = uniform(wave_max_d(on ? rx : -inf)),
                 hiy = uniform(wave_max_d(on ? ry : -inf)),
                 hiz = uniform(wave_max_d(on ? rz : -inf));
    const int n_nodes = counters[1];
    const bool any = lox <= hix;
    const double cx = uniform(0.5 * (lox + hix)), cy = uniform(0.5 * (loy + hiy)),
                 cz = uniform(0.5 * (loz + hiz));
    const float tx = on ? (float)(rx - cx) : 0.0f, ty = on ? (float)(ry - cy) : 0.0f,
                tz = on ? (float)(rz - cz) : 0.0f;
    // half extents of the target box about (cx, cy, cz), FP32, with slack
    // for the rounding of the target offsets
    const float bhx = (float)(0.5 * (hix - lox)) * 1.000001f,
                bhy = (float)(0.5 * (hiy - loy)) * 1.000001f,
                bhz = (float)(0.5 * (hiz - loz)) * 1.000001f;
    const int leaf = P.leaf_size;
    const int leaf_bits = 32 - __clz(leaf);  // bits of counts 0..leaf
    const double cs = P.crit_scale;
    const double f64 = P.f_bs;

    // one walk for all 27 copies: a stack range carries its copy index, and
    // a copy's sources are shifted by +L*shift (its targets are r - L*shift)
    double ax = 0, ay = 0, az = 0;
    unsigned long long c64 = 0, c32 = 0;
    int far_n = 0, near_n = 0;  // far: clear FP32 entries; near: FP64 entries
    const int n_copies = any ? 1 + P.n_images : 0;
    const bool recorded = g < rec_groups;  // record / replay slot of this group
    const size_t rec_base = (size_t)g * rec_cap;
    int rec_n = 0;
    bool rec_over = !recorded;

    // a zero far entry placed far away: adds exactly zero without guards
    auto pad_far = [&](int i) {
        farA[i] = make_float4(1e15f, 0.f, 0.f, 0.f);
        farB[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        farC[i] = 0.f;
    };
    // flush_far(false) evaluates whole rounds of 4 S clear entries and keeps
    // the remainder (< 4 S) at the front of the list for the next flush,
    // when the remainder and one more round of emission fit (S <= 16); the
    // group's last flush, flush_far(true), pads to whole rounds instead.
    const bool keep_rem = far_cap - (4 * S - 1) >= 64;
    auto flush_far = [&](bool last) {
        float sax = 0.f, say = 0.f, saz = 0.f, sqx = 0.f, sqy = 0.f, sqz = 0.f;
        // clear entries: whole rounds of 4 S, the last flush padded
        const bool pad = last || !keep_rem;
        const int fpad = pad ? (far_n + 4 * S - 1) / (4 * S) * (4 * S) : far_n / (4 * S) * (4 * S);
        const int nreal = pad ? far_n : fpad;  // real entries evaluated now
        for (int i = far_n + lane; i < fpad; i += 64) pad_far(i);
        __syncthreads();
        // slice s takes entries 4s .. 4s + 3 of each round of 4 S
        for (int i0 = 4 * slice; i0 < fpad; i0 += 4 * S) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + u;
                const float4 a = farA[i];
                const float4 b = farB[i];
                seg_field32_far_clear(a.x - tx, a.y - ty, a.z - tz, a.w, b.x, b.y, b.z, b.w,
                                      farC[i], sax, say, saz, sqx, sqy, sqz);
            }
        }
        // sum of m (a - t) x q = sum m a x q - t x sum m q
        ax += (double)(sax - fmaf(ty, sqz, -(tz * sqy))) * f64;
        ay += (double)(say - fmaf(tz, sqx, -(tx * sqz))) * f64;
        az += (double)(saz - fmaf(tx, sqy, -(ty * sqx))) * f64;
        if (on) {  // pairs of this slice: blocks of 4 with index = slice mod S
            const int nf = nreal >> 2, rem = nreal & 3;
            c32 += 4 * (nf > slice ? (nf - slice + S - 1) / S : 0) + (nf % S == slice ? rem : 0);
        }
        __syncthreads();
        // the remainder [fpad, far_n) moves to the front (disjoint ranges:
        // fpad >= 4 S > remainder whenever fpad > 0)
        const int keep = far_n - nreal;
        if (keep > 0 && fpad > 0) {
            if (lane < keep) {
                const float4 a = farA[fpad + lane];
                const float4 b = farB[fpad + lane];
                const float cc = farC[fpad + lane];
                farA[lane] = a;
                farB[lane] = b;
                farC[lane] = cc;
            }
            __syncthreads();
        }
        far_n = keep;
    };
    auto flush_near = [&]() {
        __syncthreads();
        // R = (source + shift) - r, Rp = R + sum ds
        auto load_near = [&](unsigned e, double &Rx, double &Ry, double &Rz, double &Qx,
                             double &Qy, double &Qz) {
            const int idx = (int)(e & kIdxMask), cp = (int)((e >> kCopyShift) & 31u);
            if (e & kNodeBit) {
                const Node &nd = nodes[idx];
                Rx = nd.cx; Ry = nd.cy; Rz = nd.cz;
                Qx = nd.qx; Qy = nd.qy; Qz = nd.qz;
            } else {
                Rx = sx[idx]; Ry = sy[idx]; Rz = sz[idx];
                Qx = sqx[idx]; Qy = sqy[idx]; Qz = sqz[idx];
            }
            if (MIR && shM[cp]) {
                const int m = shM[cp];
                double c0, c1, c2;
                const bool rv = __popc(m) & 1;
                mirror_axis(m & 1, rv, shL[cp][0], Rx, Qx, c0);
                mirror_axis(m & 2, rv, shL[cp][1], Ry, Qy, c1);
                mirror_axis(m & 4, rv, shL[cp][2], Rz, Qz, c2);
                Rx = Rx - rx;
                Ry = Ry - ry;
                Rz = Rz - rz;
                return;
            }
            Rx = (Rx + shL[cp][0]) - rx;
            Ry = (Ry + shL[cp][1]) - ry;
            Rz = (Rz + shL[cp][2]) - rz;
        };
        for (int i = slice; i < near_n; i += S) {
            double Rx, Ry, Rz, Qx, Qy, Qz, vx, vy, vz;
            load_near(nearI[i], Rx, Ry, Rz, Qx, Qy, Qz);
            // FP64 mode: every pair, so the cheaper inverse square roots pay;
            // mixed mode keeps the few near pairs' shorter code (registers)
            const bool hit =
                seg_field64_fast(Rx, Ry, Rz, Rx + Qx, Ry + Qy, Rz + Qz, f64, vx, vy, vz);
            if (hit) {
                ax += vx; ay += vy; az += vz;
            }
        }
        if (on && near_n > slice) c64 += (near_n - slice + S - 1) / S;
        __syncthreads();
        near_n = 0;
    };
    // append one entry per predicated lane: near as a packed index, far as
    // FP32 offsets from the group centre.  A far entry (offset a, vector q)
    // is classified against the target box by its Thales sphere (centre
    // a + q/2, radius h = |q|/2), inside which R.Rp < 0:
    //   clear  - the box stays outside by a margin (distance^2 >= 1.04 h^2
    //            + 2e-10 cm^2): the reference's guards (|R|, |Rp| >= 1e-8,
    //            R.Rp/|R||Rp| >= 1e-8) hold for every target, so these pairs
    //            run without them;
    //   inside - the whole box lies inside (farthest corner^2 <= 0.96 h^2):
    //            every pair fails the last guard and adds 0, as in the
    //            reference; not stored;
    //   cross  - the rest (about 2% on dense tangles: long monopoles): near
    //            entries, FP64 with the reference's guards.  The last guard
    //            is a sign test on a cosine; evaluated in FP32, a pair within
    //            ~1e-7 of the threshold could flip and add (or drop) its whole
    //            contribution, which on the 1.4M-point tangle put a few
    //            points' velocities off by percents.
    auto emit = [&](bool want, bool near, unsigned idx, float ex, float ey, float ez,
                    float gx, float gy, float gz) {
        bool clear = true, inside = false;
        if (MIXED) {
            const float mx = fabsf(fmaf(0.5f, gx, ex)), my = fabsf(fmaf(0.5f, gy, ey)),
                        mz = fabsf(fmaf(0.5f, gz, ez));
            const float dx = fmaxf(mx - bhx, 0.f), dy = fmaxf(my - bhy, 0.f),
                        dz = fmaxf(mz - bhz, 0.f);
            const float fx = mx + bhx, fy = my + bhy, fz = mz + bhz;
            const float q2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
            clear = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) >= fmaf(0.26f, q2, 2e-10f);
            inside = !clear && fmaf(fz, fz, fmaf(fy, fy, fx * fx)) <= 0.24f * q2;
        }
        const bool nr = near || !clear;  // FP64 entries: near field and cross
        const bool keep = want && !(MIXED && !near && inside);
        const unsigned long long mf = __ballot(keep && !nr);
        const unsigned long long mn = __ballot(keep && nr);
        const int cf = __popcll(mf), cn = __popcll(mn);
        if (keep && !nr) {
            const int pos = far_n + prefix_below(mf);
            farA[pos] = make_float4(ex, ey, ez, gx);
            farB[pos] = make_float4(gy, gz, fmaf(ey, gz, -(ez * gy)), fmaf(ez, gx, -(ex * gz)));
            farC[pos] = fmaf(ex, gy, -(ey * gx));
        }
        if (keep && nr) nearI[near_n + prefix_below(mn)] = idx;
        far_n += cf;
        near_n += cn;
        if (MODE == kRecord) {  // every accepted entry, whatever its class
            const unsigned long long mw = __ballot(want);
            const int cw = __popcll(mw);
            if (rec_n + cw <= rec_cap) {
                if (want) rec[rec_base + rec_n + prefix_below(mw)] = idx;
            } else {
                rec_over = true;
            }
            rec_n += cw;
        }
    };

    const int rc = (MODE == kReplay && recorded) ? rec_count[g] : -1;
    if (MODE == kReplay && rc >= 0) {
        // the recorded entries at the current positions, 64 per round
        for (int b0 = 0; b0 < rc; b0 += 64) {
            if (far_n + 64 > far_cap) flush_far(false);
            if (near_n + 64 > kListNear) flush_near();
            const int j = b0 + lane;
            const bool want = j < rc;
            const unsigned e = want ? rec[rec_base + j] : 0u;
            const int idx = (int)(e & kIdxMask), cp = (int)((e >> kCopyShift) & 31u);
            float ex = 0, ey = 0, ez = 0, gx = 0, gy = 0, gz = 0;
            double e2 = 0;
            if (want) {
                double ux, uy, uz;
                if (e & kNodeBit) {
                    const Node &nd = nodes[idx];
                    ux = nd.cx; uy = nd.cy; uz = nd.cz;
                    gx = (float)nd.qx; gy = (float)nd.qy; gz = (float)nd.qz;
                } else {
                    ux = sx[idx]; uy = sy[idx]; uz = sz[idx];
                    gx = (float)sqx[idx]; gy = (float)sqy[idx]; gz = (float)sqz[idx];
                }
                ux += shL[cp][0]; uy += shL[cp][1]; uz += shL[cp][2];
                ex = (float)(ux - cx); ey = (float)(uy - cy); ez = (float)(uz - cz);
                e2 = box_dist2(ux, uy, uz, lox, loy, loz, hix, hiy, hiz);
            }
            emit(want, !MIXED || e2 < near2, e, ex, ey, ez, gx, gy, gz);
        }
    } else {
    const int root = forest ? roots[tg] : 0;
    const int rend = forest ? nodes[root].skip : n_nodes;
    if (lane < n_copies) stk[lane] = make_int2(root, rend | (lane << kCopyShift));
    int sp = rend > root ? n_copies : 0;
    __syncthreads();
    while (sp > 0) {
        if (far_n + 64 > far_cap) flush_far(false);
        if (near_n + 64 > kListNear) flush_near();
        // each popped range pushes at most two: keep kStackMargin free
        const int room = stack_cap - kStackMargin - sp;
        const int kp = min(min(sp, 64), max(1, room));
        const bool act = lane < kp;
        // every lane reads a range and a node: an idle lane takes the stack's
        // first entry and the root of this tangle (valid addresses; its kind
        // is forced to nothing below), so the round has no per-lane branches
        const int2 rg = stk[act ? sp - kp + lane : 0];
        __syncthreads();
        sp -= kp;
        const int d = act ? rg.x : root;
        const int cp = act ? rg.y >> kCopyShift : 0, end = rg.y & (int)kIdxMask;
        const double shx = shL[cp][0], shy = shL[cp][1], shz = shL[cp][2];
        const Node nd = nodes[d];
        // the node's centre in this copy
        const double ncx = nd.cx + shx, ncy = nd.cy + shy, ncz = nd.cz + shz;
        const int skp = min(max(nd.skip, d + 1), end);  // subtree end, inside the range
        double dd2;
        if (MIR && shM[cp]) {
            const int m = shM[cp];
            dd2 = box_dist2((m & 1) ? shx - nd.cx : ncx, (m & 2) ? shy - nd.cy : ncy,
                            (m & 4) ? shz - nd.cz : ncz, lox, loy, loz, hix, hiy, hiz);
        } else {
            dd2 = box_dist2(ncx, ncy, ncz, lox, loy, loz, hix, hiy, hiz);
        }
        // 0 nothing, 1 entry, 2 bucket, 3 open
        const int kind = !act || nd.count <= 0                       ? 0
                         : (nd.count == 1 || dd2 > nd.Lm * nd.Lm * cs) ? 1
                         : (nd.count <= leaf || skp == d + 1)          ? 2
                                                                       : 3;
        int st = 0;  // first sorted segment of the cell (buckets)
        if (kind == 2) st = meta[d].start;  // only buckets read their points
        // pushes (both ranges nonempty, first index beyond d, end <= the
        // popped range's end): sibling continuation, then the children
        const bool sib = act && skp < end;
        const bool opn = kind == 3;
        const unsigned long long ms = __ballot(sib), mo = __ballot(opn);
        const int base = sp + prefix_below(ms) + prefix_below(mo);
        if (sib) stk[base] = make_int2(skp, end | (cp << kCopyShift));
        if (opn) stk[base + (int)sib] = make_int2(d + 1, skp | (cp << kCopyShift));
        sp += __popcll(ms) + __popcll(mo);
        // monopoles and single segments
        // (emit is wave-wide: the entry is formed per lane first)
        float mex = (float)(ncx - cx), mey = (float)(ncy - cy), mez = (float)(ncz - cz);
        float mgx = (float)nd.qx, mgy = (float)nd.qy, mgz = (float)nd.qz;
        if (MIR && shM[cp]) {
            const int m = shM[cp];
            double x0 = nd.cx, y0 = nd.cy, z0 = nd.cz, q0 = nd.qx, q1 = nd.qy, q2 = nd.qz;
            double c0, c1, c2;
            const bool rv = __popc(m) & 1;
            mirror_axis(m & 1, rv, shx, x0, q0, c0);
            mirror_axis(m & 2, rv, shy, y0, q1, c1);
            mirror_axis(m & 4, rv, shz, z0, q2, c2);
            mex = (float)(x0 - cx); mey = (float)(y0 - cy); mez = (float)(z0 - cz);
            mgx = (float)q0; mgy = (float)q1; mgz = (float)q2;
        }
        emit(kind == 1, !MIXED || dd2 < near2,
             (unsigned)d | ((unsigned)cp << kCopyShift) | kNodeBit, mex, mey, mez, mgx, mgy, mgz);
        // buckets: the round's bucket cells laid end to end, expanded 64
        // segments per pass (one per lane, its cell found by binary search
        // over the cells' end offsets)
        const int bc = kind == 2 ? nd.count : 0;
        int bincl, btotal;
        if (__ballot(bc > leaf)) {
            // a childless cell above leaf_size points (coincident keys)
            bincl = bc;
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_up(bincl, off, 64);
                if (lane >= off) bincl += u;
            }
            btotal = __builtin_amdgcn_readlane(bincl, 63);
        } else {
            bincl = wave_scan_small(bc, leaf_bits, btotal);
        }
        if (btotal > 0) {
            const unsigned long long mb = __ballot(bc > 0);
            const int nb = __popcll(mb);
            if (bc > 0) {
                const int slot = prefix_below(mb);
                bEnd[slot] = bincl;
                bStart[slot] = st;
                bCopy[slot] = cp;
            }
            __syncthreads();
            for (int b0 = 0; b0 < btotal; b0 += 64) {
            if (far_n + 64 > far_cap) flush_far(false);
            if (near_n + 64 > kListNear) flush_near();
                const int j = b0 + lane;
                const bool want = j < btotal;
                float ex = 0, ey = 0, ez = 0, gx = 0, gy = 0, gz = 0;
                double b2 = 0;
                int sg = 0, bcp = 0;
                if (want) {
                    int lo = 0, hi = nb - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (bEnd[mid] > j) hi = mid;
                        else lo = mid + 1;
                    }
                    sg = bStart[lo] + j - (lo ? bEnd[lo - 1] : 0);
                    bcp = bCopy[lo];
                    if (MIR && shM[bcp]) {
                        const int m = shM[bcp];
                        double x0 = sx[sg], y0 = sy[sg], z0 = sz[sg];
                        double q0 = sqx[sg], q1 = sqy[sg], q2 = sqz[sg], c0, c1, c2;
                        const bool rv = __popc(m) & 1;
                        mirror_axis(m & 1, rv, shL[bcp][0], x0, q0, c0);
                        mirror_axis(m & 2, rv, shL[bcp][1], y0, q1, c1);
                        mirror_axis(m & 4, rv, shL[bcp][2], z0, q2, c2);
                        gx = (float)q0; gy = (float)q1; gz = (float)q2;
                        ex = (float)(x0 - cx); ey = (float)(y0 - cy); ez = (float)(z0 - cz);
                        b2 = box_dist2(c0, c1, c2, lox, loy, loz, hix, hiy, hiz);
                    } else {
                    const double ux = sx[sg] + shL[bcp][0], uy = sy[sg] + shL[bcp][1],
                                 uz = sz[sg] + shL[bcp][2];
                    gx = (float)sqx[sg]; gy = (float)sqy[sg]; gz = (float)sqz[sg];
                    ex = (float)(ux - cx); ey = (float)(uy - cy); ez = (float)(uz - cz);
                    b2 = box_dist2(ux, uy, uz, lox, loy, loz, hix, hiy, hiz);
                    }
                }
                emit(want, !MIXED || b2 < near2, (unsigned)sg | ((unsigned)bcp << kCopyShift),
                     ex, ey, ez, gx, gy, gz);
            }
        }
        __syncthreads();
    }
    }  // walk
    if (MODE == kRecord && recorded && lane == 0) rec_count[g] = rec_over ? -1 : rec_n;
    if (far_n) flush_far(true);
    if (near_n) flush_near();
    // slice partials in slice order
    for (int s = 1; s < S; ++s) {
        const int src = lane % ny + s * ny;
        const double bx = __shfl(ax, src, 64), by = __shfl(ay, src, 64),
                     bz = __shfl(az, src, 64);
        if (slice == 0) {
            ax += bx; ay += by; az += bz;
        }
    }
    if (valid && slice == 0) {
        // all copies in one sum: the primary/image split of the reference's
        // summation order is not kept by this traversal
        const double nan = __builtin_nan("");
        px[k] = on ? ax : nan; py[k] = on ? ay : nan; pz[k] = on ? az : nan;
        ix[k] = on ? 0.0 : nan; iy[k] = on ? 0.0 : nan; iz[k] = on ? 0.0 : nan;
    }
    for (int off = 32; off > 0; off >>= 1) {
        c64 += __shfl_xor(c64, off, 64);
        c32 += __shfl_xor(c32, off, 64);
    }
    if (lane == 0) {
        if (c64) atomicAdd(icount, c64);
        if (c32) atomicAdd(icount + 1, c32);
    }
    __syncthreads();
    }
}

// ---------------------------------------------------------- target groups --
// Groups are runs of at most PPW consecutive sorted points that never cross
// the boundary of an octree cell of level L0, so each group's bounding box
// is at most one such cell.  cut(i): point i starts a level-L0 cell
// (common levels of sorted keys i-1 and i below L0).
// In a forest each tangle has its own level (from its key); the boundary
// between two tangles is always a cut (their keys share no level).
__device__ __forceinline__ int cell_cut(const int8_t *__restrict__ delta,
                                        const uint64_t *__restrict__ keys,
                                        const TangleParams *__restrict__ Pp, int i, int L0)
{
    if (i == 0) return 1;
    if (Pp[0].n_tangles > 1) L0 = Pp[key_tangle(keys[i])].group_level;
    return delta[i - 1] < L0;
}

__global__ void k_grp_cut(int n, const int8_t *__restrict__ delta,
                          const uint64_t *__restrict__ keys, const TangleParams *__restrict__ Pp,
                          int L0, int *__restrict__ a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = cell_cut(delta, keys, Pp, i, L0);
}

// a = exclusive scan of cuts: cell index b[i]; c[cell] = first point,
// c[n_cells] = n.
__global__ void k_grp_cells(int n, const int8_t *__restrict__ delta,
                            const uint64_t *__restrict__ keys,
                            const TangleParams *__restrict__ Pp, int L0,
                            const int *__restrict__ a, int *__restrict__ b, int *__restrict__ c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cut = cell_cut(delta, keys, Pp, i, L0);
    int cell = a[i] + cut - 1;
    b[i] = cell;
    if (cut) c[cell] = i;
    if (i == n - 1) c[a[n]] = n;
}

template <int PPW>
__global__ void k_grp_flag(int n, const int *__restrict__ b, const int *__restrict__ c,
                           int *__restrict__ a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (i - c[b[i]]) % PPW == 0;
}

// a = exclusive scan of group starts
template <int PPW>
__global__ void k_grp_emit(int n, const int *__restrict__ a, const int *__restrict__ b,
                           const int *__restrict__ c, int2 *__restrict__ groups,
                           int *__restrict__ counters)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cell = b[i];
    if ((i - c[cell]) % PPW == 0) groups[a[i]] = make_int2(i, min(PPW, c[cell + 1] - i));
    if (i == 0) counters[kGroupCounter] = a[n];
}

int list_group_level(const ovg_params &p)
{
    if (p.group_level > 0) return p.group_level;
    if (!(p.mixed_near > 0)) return 5;
    double side = 0;
    for (int k = 0; k < 3; ++k) side = std::max(side, p.box_hi[k] - p.box_lo[k]);
    int L = 1;
    while (L < 10 && side / (double)(1 << L) > 8 * p.mixed_near) ++L;
    return L;
}

template <int PPW>
static int build_groups(Ctx &c)
{
    TreeBufs &T = c.tree;
    const int n = c.n_live, L0 = list_group_level(c.params);
    OVG_TRY(T.grp_a.ensure(sizeof(int) * (n + 1)));
    OVG_TRY(T.grp_b.ensure(sizeof(int) * (n + 1)));
    OVG_TRY(T.grp_c.ensure(sizeof(int) * (n + 2)));
    OVG_TRY(T.groups.ensure(sizeof(int2) * (n + 1)));
    int *a = T.grp_a.as<int>(), *b = T.grp_b.as<int>(), *cc = T.grp_c.as<int>();
    const int8_t *delta = T.delta.as<int8_t>();
    const int gs = (n + 255) / 256;
    const uint64_t *keys = T.keys[0].as<uint64_t>();
    hipLaunchKernelGGL(k_grp_cut, dim3(gs), dim3(256), 0, c.stream, n, delta, keys, dparams(c),
                       L0, a);
    OVG_TRY(launch_exclusive_scan(c, a, a, n, a + n, T.scan_tmp));
    hipLaunchKernelGGL(k_grp_cells, dim3(gs), dim3(256), 0, c.stream, n, delta, keys,
                       dparams(c), L0, a, b, cc);
    hipLaunchKernelGGL(k_grp_flag<PPW>, dim3(gs), dim3(256), 0, c.stream, n, b, cc, a);
    OVG_TRY(launch_exclusive_scan(c, a, a, n, a + n, T.scan_tmp));
    hipLaunchKernelGGL(k_grp_emit<PPW>, dim3(gs), dim3(256), 0, c.stream, n, a, b, cc,
                       T.groups.as<int2>(), T.counters.as<int>());
    return OVG_OK;
}

int list_points_per_wave(const ovg_params &p)
{
    return p.points_per_wave == 0 ? 16 : p.points_per_wave;
}

// The launch shape (validated in ovg_set_params: stack 128..384, tile 96 or
// 128, grid 0..64 wavefronts per CU).
static int list_stack_cap(const ovg_params &p)
{
    return p.list_stack == 0 ? kListStack : std::min(p.list_stack, kListStack);
}

static int list_far_tile(const ovg_params &p)
{
    return p.list_tile == 96 ? 96 : kListFar;
}

static int list_grid(const Ctx &c, int groups)
{
    const int b = c.params.list_blocks;
    return b > 0 ? std::max(1, std::min(groups, b * c.n_cu)) : groups;
}

int launch_bh_list(Ctx &c, const double *const x[3])
{
    if (c.n_live == 0) return OVG_OK;
    if (c.n_live >= (1 << 25))
        return set_error(OVG_ELIMIT, "list traversal: %d points exceed its 2^25 limit",
                         c.n_live);
    TreeBufs &T = c.tree;
    const int ppw = list_points_per_wave(c.params);
    int mode = c.list_mode;
    bool mirrored = false;
    for (int j = 0; j < c.params.n_images; ++j) mirrored |= c.params.image_reflect[j] != 0;
    for (const ovg_params &q : c.forest_params)
        for (int j = 0; j < q.n_images; ++j) mirrored |= q.image_reflect[j] != 0;
    if (mirrored) mode = kMirror;  // reuse_lists is refused with mirror images
    // a replay keeps the recorded run's groups (same tree, same order)
    if (mode != kReplay) {
        switch (ppw) {
        case 8: OVG_TRY(build_groups<8>(c)); break;
        case 16: OVG_TRY(build_groups<16>(c)); break;
        case 21: OVG_TRY(build_groups<21>(c)); break;
        case 32: OVG_TRY(build_groups<32>(c)); break;
        default: OVG_TRY(build_groups<64>(c)); break;
        }
    }
    // groups are counted on the device; blocks stride over them, so any
    // count is covered (typically 1.0-1.2 x n / ppw).  list_blocks > 0: a
    // persistent grid of that many wavefronts per CU.
    const int full = (c.n_live + ppw - 1) / ppw;
    const int slots = full + full / 4 + 8;
    const int blocks = list_grid(c, slots);
    // records: one slot of kRecCap entries for each of the first `slots`
    // groups (a group beyond them, or one whose entries do not fit, walks)
    int rec_groups = 0, rec_cap = 0;
    if (mode == kRecord) {
        OVG_TRY(T.rec.ensure(sizeof(unsigned) * (size_t)slots * kRecCap));
        OVG_TRY(T.rec_count.ensure(sizeof(int) * (size_t)slots));
        rec_groups = slots;
        rec_cap = kRecCap;
        c.rec_groups = slots;
    } else if (mode == kReplay) {
        rec_groups = c.rec_groups;
        rec_cap = kRecCap;
    }
    unsigned long long *cnt = c.counter64.as<unsigned long long>();
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64), 0, c.stream, T.rec.as<unsigned>(),
                           T.rec_count.as<int>(), rec_cap, rec_groups, T.groups.as<int2>(),
                           T.sorted_pts.as<int>(), x[0], x[1], x[2], T.nodes.as<Node>(),
                           T.meta.as<NodeMeta>(), T.sx.as<double>(), T.sy.as<double>(),
                           T.sz.as<double>(), T.sqx.as<double>(), T.sqy.as<double>(),
                           T.sqz.as<double>(), T.counters.as<int>(),
                           T.counters.as<int>() + kGroupCounter, dparams(c),
                           T.roots.as<int>(), c.bs_prim[0].as<double>(), c.bs_prim[1].as<double>(),
                           c.bs_prim[2].as<double>(), c.bs_img[0].as<double>(),
                           c.bs_img[1].as<double>(), c.bs_img[2].as<double>(), cnt,
                           list_stack_cap(c.params), list_far_tile(c.params));
    };
    const bool mixed = c.params.precision == OVG_PREC_MIXED;
    auto by_mode = [&](auto kw, auto kr, auto kp, auto km) {
        if (mode == kRecord) go(kr);
        else if (mode == kReplay) go(kp);
        else if (mode == kMirror) go(km);
        else go(kw);
    };
#define OVG_LIST_MODES(M, W)                                                                \
    by_mode(k_bh_list<M, W, kWalk>, k_bh_list<M, W, kRecord>, k_bh_list<M, W, kReplay>, \
            k_bh_list<M, W, kMirror>)
    switch (ppw) {
    case 8: mixed ? OVG_LIST_MODES(true, 8) : OVG_LIST_MODES(false, 8); break;
    case 16: mixed ? OVG_LIST_MODES(true, 16) : OVG_LIST_MODES(false, 16); break;
    case 21: mixed ? OVG_LIST_MODES(true, 21) : OVG_LIST_MODES(false, 21); break;
    case 32: mixed ? OVG_LIST_MODES(true, 32) : OVG_LIST_MODES(false, 32); break;
    default: mixed ? OVG_LIST_MODES(true, 64) : OVG_LIST_MODES(false, 64); break;
    }
#undef OVG_LIST_MODES
    return OVG_OK;
}

// ----------------------------------------------------- target subsets --
__global__ void k_rank_key(int m, const int *__restrict__ slots, const int *__restrict__ rank,
                           const uint64_t *__restrict__ keys, int *__restrict__ out_rank,
                           uint64_t *__restrict__ out_key)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int r = rank[slots[i]];
    out_rank[i] = r;
    out_key[i] = keys[r];
}

// Same levels of 3-bit digits as k_tree.hip's LCP levels of tuned keys
// (one tangle: the Morton digits of 16 levels at key bits 1..48).
static int host_lcp_levels(uint64_t a, uint64_t b)
{
    if (a == b) return 16;
    return (__builtin_clzll(a ^ b) - 15) / 3;
}

int launch_bh_list_targets(Ctx &c, const double *const x[3], int m, const int *d_slots)
{
    if (m <= 0) return OVG_OK;
    TreeBufs &T = c.tree;
    const int ppw = list_points_per_wave(c.params), L0 = list_group_level(c.params);
    // tree order and keys of the targets, grouped on the host
    OVG_TRY(T.sub_tgt.ensure((sizeof(int) + sizeof(uint64_t)) * (size_t)m + 64));
    int *d_rank = T.sub_tgt.as<int>();
    uint64_t *d_key = reinterpret_cast<uint64_t *>(
        T.sub_tgt.as<unsigned char>() + ((sizeof(int) * (size_t)m + 15) & ~(size_t)15));
    hipLaunchKernelGGL(k_rank_key, dim3((m + 255) / 256), dim3(256), 0, c.stream, m, d_slots,
                       T.rank.as<int>(), T.keys[0].as<uint64_t>(), d_rank, d_key);
    std::vector<int> rk(m), sl(m);
    std::vector<uint64_t> ky(m);
    OVG_HIP(hipMemcpyAsync(rk.data(), d_rank, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost,
                           c.stream));
    OVG_HIP(hipMemcpyAsync(ky.data(), d_key, sizeof(uint64_t) * (size_t)m,
                           hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipMemcpyAsync(sl.data(), d_slots, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost,
                           c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    std::vector<int> order(m);
    for (int i = 0; i < m; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return rk[a] < rk[b]; });
    std::vector<int> tgt(m);
    std::vector<int2> grp;
    for (int i = 0; i < m; ++i) {
        tgt[i] = sl[order[i]];
        const bool cut = i == 0 || grp.back().y == ppw ||
                         host_lcp_levels(ky[order[i - 1]], ky[order[i]]) < L0;
        if (cut) grp.push_back(make_int2(i, 1));
        else grp.back().y++;
    }
    const int ng = (int)grp.size();
    OVG_TRY(T.sub_groups.ensure(sizeof(int2) * (size_t)ng + sizeof(int) * (size_t)m + 64));
    int2 *d_grp = T.sub_groups.as<int2>();
    int *d_tgt = reinterpret_cast<int *>(d_grp + ng);
    OVG_TRY(T.sub_count.ensure(sizeof(int)));
    OVG_HIP(hipMemcpyAsync(d_grp, grp.data(), sizeof(int2) * (size_t)ng, hipMemcpyHostToDevice,
                           c.stream));
    OVG_HIP(hipMemcpyAsync(d_tgt, tgt.data(), sizeof(int) * (size_t)m, hipMemcpyHostToDevice,
                           c.stream));
    OVG_HIP(hipMemcpyAsync(T.sub_count.p, &ng, sizeof(int), hipMemcpyHostToDevice, c.stream));
    OVG_TRY(launch_bh_list_groups(c, x, d_grp, d_tgt, T.sub_count.as<int>(), ng));
    // the host vectors must outlive the copies above
    OVG_HIP(hipStreamSynchronize(c.stream));
    return OVG_OK;
}

int launch_bh_list_groups(Ctx &c, const double *const x[3], const int2 *d_grp, const int *d_tgt,
                          const int *d_ngroups, int max_groups)
{
    if (max_groups <= 0) return OVG_OK;
    TreeBufs &T = c.tree;
    const int ppw = list_points_per_wave(c.params);
    // the count may be a capacity bound (groups counted on the device):
    // blocks stride over the groups, at most 16 wavefronts per CU
    const int grid = std::min(list_grid(c, max_groups), 16 * c.n_cu);
    unsigned long long *cnt = c.counter64.as<unsigned long long>();
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, c.stream, nullptr,
                           nullptr, 0, 0, d_grp, d_tgt, x[0], x[1],
                           x[2], T.nodes.as<Node>(), T.meta.as<NodeMeta>(), T.sx.as<double>(),
                           T.sy.as<double>(), T.sz.as<double>(), T.sqx.as<double>(),
                           T.sqy.as<double>(), T.sqz.as<double>(), T.counters.as<int>(),
                           d_ngroups, dparams(c), T.roots.as<int>(),
                           c.bs_prim[0].as<double>(), c.bs_prim[1].as<double>(),
                           c.bs_prim[2].as<double>(), c.bs_img[0].as<double>(),
                           c.bs_img[1].as<double>(), c.bs_img[2].as<double>(), cnt,
                           list_stack_cap(c.params), list_far_tile(c.params));
    };
    const bool mixed = c.params.precision == OVG_PREC_MIXED;
    bool mirrored = false;
    for (int j = 0; j < c.params.n_images; ++j) mirrored |= c.params.image_reflect[j] != 0;
#define OVG_LIST_WALK(M, W) \
    (mirrored ? go(k_bh_list<M, W, kMirror>) : go(k_bh_list<M, W, kWalk>))
    switch (ppw) {
    case 8: mixed ? OVG_LIST_WALK(true, 8) : OVG_LIST_WALK(false, 8); break;
    case 16: mixed ? OVG_LIST_WALK(true, 16) : OVG_LIST_WALK(false, 16); break;
    case 21: mixed ? OVG_LIST_WALK(true, 21) : OVG_LIST_WALK(false, 21); break;
    case 32: mixed ? OVG_LIST_WALK(true, 32) : OVG_LIST_WALK(false, 32); break;
    default: mixed ? OVG_LIST_WALK(true, 64) : OVG_LIST_WALK(false, 64); break;
    }
#undef OVG_LIST_WALK
    return launch_status();
}

}  // namespace ovg
