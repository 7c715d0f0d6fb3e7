// Barnes-Hut traversal (octree_get_vs, src/octree.c:296-356) over the
// DFS-ordered compressed octree of k_tree.hip.  The recursion becomes a
// stackless walk: at node d a point either drops the subtree (distance to
// the centre of mass < 1e-8, :316), accepts the node's monopole (Lm/d <
// theta or N == 1, :319) and jumps to d's skip rope, or descends to d+1.
//
// k_bh_point: one lane per target point, images in sequence, the
//   reference's acceptance test Lm/d < theta in FP64 with d = |r - com|
//   correctly rounded.  This is the plain FP64 path.
// k_bh_group<.., LOCKSTEP = true>: one wavefront per 64 Morton-adjacent
//   target points and all images.  Each lane keeps its own next node; the wave visits the minimum
//   over its lanes, loading that node once (scalar, 64 B) and letting only
//   the lanes positioned there act on it.  Every lane therefore sees exactly
//   the node sequence of its own per-point walk.  Acceptance uses
//   |r - com|^2 > (Lm/theta)^2, equal to the reference test except at exact
//   ties.  Nodes with count <= leaf_size that are not accepted are summed
//   segment by segment (buckets).  With OVG_PREC_MIXED internal-node
//   monopoles and far segments run in FP32 (R formed in FP64, then rounded),
//   segments nearer than mixed_near in FP64; per-point sums are FP64.
// k_bh_group<.., LOCKSTEP = false>: the same decisions with every lane
//   walking on its own, for scattered target lists (reconnection endpoints).
// VALU-bound in the pair formula, latency-bound in the node walk.
#include "engine_internal.h"

namespace ovg {

constexpr int kBhBlock = 256;

__global__ __launch_bounds__(kBhBlock) void k_bh_point(
    int n_live, const int *__restrict__ spts, const int *__restrict__ tlist,
    const double *__restrict__ X,
    const double *__restrict__ Y, const double *__restrict__ Z, const Node *__restrict__ nodes,
    const int *__restrict__ counters, const TangleParams *__restrict__ Pp,
    double *__restrict__ px, double *__restrict__ py, double *__restrict__ pz,
    double *__restrict__ ix, double *__restrict__ iy, double *__restrict__ iz,
    unsigned long long *__restrict__ icount)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    const TangleParams &P = *Pp;
    unsigned long long cnt = 0;
    if (j < n_live) {
        int n_nodes = counters[1];
        int k = tlist ? tlist[j] : spts[j];
        double rx = X[k], ry = Y[k], rz = Z[k];
        double tx = 0, ty = 0, tz = 0;
        // a non-finite point would open every node; its velocity is NaN
        const bool finite = isfinite(rx) && isfinite(ry) && isfinite(rz);
        if (!finite) n_nodes = 0, tx = ty = tz = __builtin_nan("");
        for (int im = -1; im < P.n_images; ++im) {
            double qx = rx, qy = ry, qz = rz;
            if (im >= 0) image_point(P, im, rx, ry, rz, qx, qy, qz);
            double ax = finite ? 0 : __builtin_nan(""), ay = ax, az = ax;
            int d = 0;
            while (d < n_nodes) {
                const Node nd = nodes[d];
                if (nd.count == 0) { d = nd.skip; continue; }
                double ex = qx - nd.cx, ey = qy - nd.cy, ez = qz - nd.cz;
                double dist = sqrt(dot3(ex, ey, ez, ex, ey, ez));
                if (dist < 1e-8) { d = nd.skip; continue; }
                if (nd.Lm / dist < P.theta || nd.count == 1) {
                    double Rx = nd.cx - qx, Ry = nd.cy - qy, Rz = nd.cz - qz;
                    double vx, vy, vz;
                    ++cnt;
                    if (seg_field64(Rx, Ry, Rz, Rx + nd.qx, Ry + nd.qy, Rz + nd.qz, P.f_bs, vx,
                                    vy, vz)) {
                        ax += vx; ay += vy; az += vz;
                    }
                    d = nd.skip;
                } else {
                    d = d + 1;
                }
            }
            if (im < 0) {
                px[k] = ax; py[k] = ay; pz[k] = az;
            } else {
                image_reflect(P, im, ax, ay, az);  // src/tangle.c:644-646
                tx += ax; ty += ay; tz += az;
            }
        }
        ix[k] = tx; iy[k] = ty; iz[k] = tz;
    }
    // block-reduce the interaction count
    __shared__ unsigned long long red[kBhBlock / 64];
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < kBhBlock / 64; ++w) s += red[w];
        if (s) atomicAdd(icount, s);
    }
}

__device__ __forceinline__ int wave_min(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

// Read-only view of the tree for the walks below.
struct TreeView {
    const Node *nodes;
    const NodeMeta *meta;
    const double *sx, *sy, *sz, *sqx, *sqy, *sqz;
    int n_nodes, leaf;
    double cs, f64, near2;
    float f32;
};

// One target copy q against the tree, adding into (ax, ay, az).  Each lane
// sees the node sequence of the reference's per-point walk.  LOCKSTEP: the
// wave visits min(p) over its lanes (coherent Morton-ordered targets);
// otherwise every lane walks alone.  on == false: the lane takes no part.
template <bool MIXED, bool LOCKSTEP>
__device__ __forceinline__ void tree_walk(const TreeView &T, bool on, double qx, double qy,
                                          double qz, double &ax, double &ay, double &az,
                                          unsigned long long &c64, unsigned long long &c32)
{
    int p = on ? 0 : 0x7fffffff;
    for (;;) {
        int d;
        if (LOCKSTEP) {
            d = __builtin_amdgcn_readfirstlane(wave_min(p));
            if (d >= T.n_nodes) break;
        } else {
            d = p;
            if (d >= T.n_nodes) break;
        }
        const Node nd = T.nodes[d];
        if (LOCKSTEP && p != d) continue;
        if (nd.count == 0) { p = nd.skip; continue; }
        double ex = nd.cx - qx, ey = nd.cy - qy, ez = nd.cz - qz;
        double d2 = ex * ex + ey * ey + ez * ez;
        if (d2 < 1e-16) { p = nd.skip; continue; }
        bool leafnode = nd.count == 1;
        if (leafnode || d2 > nd.Lm * nd.Lm * T.cs) {
            // monopole / single segment: R = com - r, Rp = R + sum ds
            if (MIXED && !(leafnode && d2 < T.near2)) {
                if (seg_field_mixed(ex, ey, ez, ex + nd.qx, ey + nd.qy, ez + nd.qz, T.f64,
                                    T.f32, ax, ay, az))
                    ++c32;
                else
                    ++c64;
            } else {
                double vx, vy, vz;
                if (seg_field64(ex, ey, ez, ex + nd.qx, ey + nd.qy, ez + nd.qz, T.f64, vx, vy,
                                vz)) {
                    ax += vx; ay += vy; az += vz;
                }
                ++c64;
            }
            p = nd.skip;
        } else if (nd.count <= T.leaf) {
            // bucket: every segment of the cell, as count-1 leaves would be
            int s0 = T.meta[d].start, s1 = s0 + nd.count;
            for (int s = s0; s < s1; ++s) {
                double bx = T.sx[s] - qx, by = T.sy[s] - qy, bz = T.sz[s] - qz;
                double b2 = bx * bx + by * by + bz * bz;
                if (b2 < 1e-16) continue;
                if (MIXED && !(b2 < T.near2)) {
                    if (seg_field_mixed(bx, by, bz, bx + T.sqx[s], by + T.sqy[s],
                                        bz + T.sqz[s], T.f64, T.f32, ax, ay, az))
                        ++c32;
                    else
                        ++c64;
                } else {
                    double vx, vy, vz;
                    if (seg_field64(bx, by, bz, bx + T.sqx[s], by + T.sqy[s], bz + T.sqz[s],
                                    T.f64, vx, vy, vz)) {
                        ax += vx; ay += vy; az += vz;
                    }
                    ++c64;
                }
            }
            p = nd.skip;
        } else {
            p = d + 1;
        }
    }
}

__device__ __forceinline__ TreeView tree_view(const Node *nodes, const NodeMeta *meta,
                                              const double *sx, const double *sy,
                                              const double *sz, const double *sqx,
                                              const double *sqy, const double *sqz,
                                              const int *counters, const TangleParams &P,
                                              double near2)
{
    TreeView T;
    T.nodes = nodes; T.meta = meta;
    T.sx = sx; T.sy = sy; T.sz = sz; T.sqx = sqx; T.sqy = sqy; T.sqz = sqz;
    T.n_nodes = counters[1];
    T.leaf = P.leaf_size;
    T.cs = P.crit_scale;
    T.f64 = P.f_bs;
    T.f32 = (float)P.f_bs;
    T.near2 = near2;
    return T;
}

// Block-sum the two interaction counters into icount[0..1].
__device__ __forceinline__ void add_counts(unsigned long long c64, unsigned long long c32,
                                           unsigned long long *icount)
{
    __shared__ unsigned long long red[2][kBhBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        c64 += __shfl_xor(c64, off, 64);
        c32 += __shfl_xor(c32, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = c64;
        red[1][threadIdx.x >> 6] = c32;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s64 = 0, s32 = 0;
        for (int w = 0; w < kBhBlock / 64; ++w) {
            s64 += red[0][w];
            s32 += red[1][w];
        }
        if (s64) atomicAdd(icount, s64);
        if (s32) atomicAdd(icount + 1, s32);
    }
}

// One lane per target, all 27 copies in sequence: primary into p*, the sum
// of the 26 images (in table order) into i*.
template <bool MIXED, bool LOCKSTEP>
__global__ __launch_bounds__(kBhBlock) void k_bh_group(
    int n_live, const int *__restrict__ spts, const int *__restrict__ tlist,
    const double *__restrict__ X,
    const double *__restrict__ Y, const double *__restrict__ Z, const Node *__restrict__ nodes,
    const NodeMeta *__restrict__ meta, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ sqx, const double *__restrict__ sqy,
    const double *__restrict__ sqz, const int *__restrict__ counters,
    const TangleParams *__restrict__ Pp, double near2, double *__restrict__ px,
    double *__restrict__ py, double *__restrict__ pz, double *__restrict__ ix,
    double *__restrict__ iy, double *__restrict__ iz, unsigned long long *__restrict__ icount)
{
    const TangleParams &P = *Pp;
    const TreeView T = tree_view(nodes, meta, sx, sy, sz, sqx, sqy, sqz, counters, P, near2);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = j < n_live;
    unsigned long long c64 = 0, c32 = 0;
    int k = valid ? (tlist ? tlist[j] : spts[j]) : 0;
    double rx = valid ? X[k] : 0, ry = valid ? Y[k] : 0, rz = valid ? Z[k] : 0;
    // a non-finite point would open every node; it takes no part and gets NaN
    const bool walk = valid && isfinite(rx) && isfinite(ry) && isfinite(rz);
    const double nan0 = walk ? 0.0 : __builtin_nan("");
    double tx = nan0, ty = nan0, tz = nan0;
    for (int im = -1; im < P.n_images; ++im) {
        double qx = rx, qy = ry, qz = rz;
        if (im >= 0) image_point(P, im, rx, ry, rz, qx, qy, qz);
        double ax = nan0, ay = nan0, az = nan0;
        tree_walk<MIXED, LOCKSTEP>(T, walk, qx, qy, qz, ax, ay, az, c64, c32);
        if (valid) {
            if (im < 0) {
                px[k] = ax; py[k] = ay; pz[k] = az;
            } else {
                image_reflect(P, im, ax, ay, az);  // src/tangle.c:644-646
                tx += ax; ty += ay; tz += az;
            }
        }
    }
    if (valid) {
        ix[k] = tx; iy[k] = ty; iz[k] = tz;
    }
    add_counts(c64, c32, icount);
}

// Scattered target lists (reconnection endpoints): one lane per (copy,
// target), lane j = c * n_tgt + t with copy c = 0 the primary and c = 1..26
// the images, each walking alone; part[c][t] holds the copy's sum.
template <bool MIXED>
__global__ __launch_bounds__(kBhBlock) void k_bh_split(
    int n_tgt, int n_copies, const int *__restrict__ tlist, const double *__restrict__ X,
    const double *__restrict__ Y, const double *__restrict__ Z, const Node *__restrict__ nodes,
    const NodeMeta *__restrict__ meta, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ sqx, const double *__restrict__ sqy,
    const double *__restrict__ sqz, const int *__restrict__ counters,
    const TangleParams *__restrict__ Pp, double near2, double *__restrict__ part,
    unsigned long long *__restrict__ icount)
{
    const TangleParams &P = *Pp;
    const TreeView T = tree_view(nodes, meta, sx, sy, sz, sqx, sqy, sqz, counters, P, near2);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = j < n_tgt * n_copies;
    unsigned long long c64 = 0, c32 = 0;
    int c = valid ? j / n_tgt : 0, t = valid ? j - c * n_tgt : 0;
    int k = valid ? tlist[t] : 0;
    double qx = valid ? X[k] : 0, qy = valid ? Y[k] : 0, qz = valid ? Z[k] : 0;
    const bool walk = valid && isfinite(qx) && isfinite(qy) && isfinite(qz);
    if (c > 0) image_point(P, c - 1, qx, qy, qz, qx, qy, qz);
    const double nan0 = walk ? 0.0 : __builtin_nan("");
    double ax = nan0, ay = nan0, az = nan0;
    tree_walk<MIXED, false>(T, walk, qx, qy, qz, ax, ay, az, c64, c32);
    if (c > 0) image_reflect(P, c - 1, ax, ay, az);  // src/tangle.c:644-646
    if (valid) {
        size_t o = (size_t)3 * j;
        part[o] = ax; part[o + 1] = ay; part[o + 2] = az;
    }
    add_counts(c64, c32, icount);
}

// Primary = part[0][t]; images = 0 + part[1][t] + ... + part[26][t], the
// order of the sequential loop in k_bh_group.
__global__ void k_bh_split_sum(int n_tgt, int n_copies, const int *__restrict__ tlist,
                               const double *__restrict__ part, double *__restrict__ px,
                               double *__restrict__ py, double *__restrict__ pz,
                               double *__restrict__ ix, double *__restrict__ iy,
                               double *__restrict__ iz)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tgt) return;
    int k = tlist[t];
    const double *q = part + (size_t)3 * t;
    px[k] = q[0]; py[k] = q[1]; pz[k] = q[2];
    double nan0 = (q[0] == q[0]) ? 0.0 : __builtin_nan("");
    double tx = nan0, ty = nan0, tz = nan0;
    for (int c = 1; c < n_copies; ++c) {
        const double *r = part + (size_t)3 * ((size_t)c * n_tgt + t);
        tx += r[0]; ty += r[1]; tz += r[2];
    }
    ix[k] = tx; iy[k] = ty; iz[k] = tz;
}

// Traversal for every live point (tlist == nullptr, targets in Morton
// order) or for the n_tgt slots of tlist.
void launch_bh_targets(Ctx &c, const double *const x[3], int n_tgt, const int *tlist)
{
    if (n_tgt == 0) return;
    TreeBufs &T = c.tree;
    int gs = (n_tgt + kBhBlock - 1) / kBhBlock;
    unsigned long long *cnt = c.counter64.as<unsigned long long>();
    if (c.params.traverse == OVG_TRAVERSE_POINT) {
        hipLaunchKernelGGL(k_bh_point, dim3(gs), dim3(kBhBlock), 0, c.stream, n_tgt,
                           T.sorted_pts.as<int>(), tlist, x[0], x[1], x[2], T.nodes.as<Node>(),
                           T.counters.as<int>(), dparams(c), c.bs_prim[0].as<double>(),
                           c.bs_prim[1].as<double>(), c.bs_prim[2].as<double>(),
                           c.bs_img[0].as<double>(), c.bs_img[1].as<double>(),
                           c.bs_img[2].as<double>(), cnt);
        return;
    }
    double near2 = c.params.mixed_near * c.params.mixed_near;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(gs), dim3(kBhBlock), 0, c.stream, n_tgt,
                           T.sorted_pts.as<int>(), tlist, x[0], x[1], x[2], T.nodes.as<Node>(),
                           T.meta.as<NodeMeta>(), T.sx.as<double>(), T.sy.as<double>(),
                           T.sz.as<double>(), T.sqx.as<double>(), T.sqy.as<double>(),
                           T.sqz.as<double>(), T.counters.as<int>(), dparams(c), near2,
                           c.bs_prim[0].as<double>(), c.bs_prim[1].as<double>(),
                           c.bs_prim[2].as<double>(), c.bs_img[0].as<double>(),
                           c.bs_img[1].as<double>(), c.bs_img[2].as<double>(), cnt);
    };
    bool mixed = c.params.precision == OVG_PREC_MIXED;
    if (tlist) {
        int n_copies = c.params.n_images + 1;
        size_t lanes = (size_t)n_tgt * n_copies;
        if (lanes > 0x7fffffff) {
            // beyond int lane indices: one lane per target
            if (mixed) go(k_bh_group<true, false>);
            else go(k_bh_group<false, false>);
            return;
        }
        if (c.split_part.ensure(lanes * 3 * sizeof(double)) != OVG_OK) {
            if (mixed) go(k_bh_group<true, false>);
            else go(k_bh_group<false, false>);
            return;
        }
        int gl = (int)((lanes + kBhBlock - 1) / kBhBlock);
        auto split = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(gl), dim3(kBhBlock), 0, c.stream, n_tgt, n_copies,
                               tlist, x[0], x[1], x[2], T.nodes.as<Node>(),
                               T.meta.as<NodeMeta>(), T.sx.as<double>(), T.sy.as<double>(),
                               T.sz.as<double>(), T.sqx.as<double>(), T.sqy.as<double>(),
                               T.sqz.as<double>(), T.counters.as<int>(), dparams(c), near2,
                               c.split_part.as<double>(), cnt);
        };
        if (mixed) split(k_bh_split<true>);
        else split(k_bh_split<false>);
        hipLaunchKernelGGL(k_bh_split_sum, dim3((n_tgt + 255) / 256), dim3(256), 0, c.stream,
                           n_tgt, n_copies, tlist, c.split_part.as<double>(),
                           c.bs_prim[0].as<double>(), c.bs_prim[1].as<double>(),
                           c.bs_prim[2].as<double>(), c.bs_img[0].as<double>(),
                           c.bs_img[1].as<double>(), c.bs_img[2].as<double>());
    } else {
        if (mixed) go(k_bh_group<true, true>);
        else go(k_bh_group<false, true>);
    }
}

int launch_bh(Ctx &c, const double *const x[3])
{
    if (c.params.traverse == OVG_TRAVERSE_LIST) return launch_bh_list(c, x);
    launch_bh_targets(c, x, c.n_live, nullptr);
    return OVG_OK;
}

}  // namespace ovg
