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
            if (im >= 0) {
                qx = rx - P.Ls[im][0];
                qy = ry - P.Ls[im][1];
                qz = rz - P.Ls[im][2];
            }
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

// LOCKSTEP: the wave visits min(p) over its lanes (coherent Morton-ordered
// targets); otherwise every lane walks alone (scattered target lists).
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
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = j < n_live;
    const int n_nodes = counters[1];
    const int leaf = P.leaf_size;
    const double cs = P.crit_scale;
    const double f64 = P.f_bs;
    const float f32 = (float)P.f_bs;
    unsigned long long c64 = 0, c32 = 0;
    int k = valid ? (tlist ? tlist[j] : spts[j]) : 0;
    double rx = valid ? X[k] : 0, ry = valid ? Y[k] : 0, rz = valid ? Z[k] : 0;
    // a non-finite point would open every node; it takes no part and gets NaN
    const bool walk = valid && isfinite(rx) && isfinite(ry) && isfinite(rz);
    const double nan0 = walk ? 0.0 : __builtin_nan("");
    double tx = nan0, ty = nan0, tz = nan0;
    for (int im = -1; im < P.n_images; ++im) {
        double qx = rx, qy = ry, qz = rz;
        if (im >= 0) {
            qx = rx - P.Ls[im][0];
            qy = ry - P.Ls[im][1];
            qz = rz - P.Ls[im][2];
        }
        double ax = nan0, ay = nan0, az = nan0;
        int p = walk ? 0 : 0x7fffffff;
        for (;;) {
            int d;
            if (LOCKSTEP) {
                d = __builtin_amdgcn_readfirstlane(wave_min(p));
                if (d >= n_nodes) break;
            } else {
                d = p;
                if (d >= n_nodes) break;
            }
            const Node nd = nodes[d];
            if (LOCKSTEP && p != d) continue;
            if (nd.count == 0) { p = nd.skip; continue; }
            double ex = nd.cx - qx, ey = nd.cy - qy, ez = nd.cz - qz;
            double d2 = ex * ex + ey * ey + ez * ez;
            if (d2 < 1e-16) { p = nd.skip; continue; }
            bool leafnode = nd.count == 1;
            if (leafnode || d2 > nd.Lm * nd.Lm * cs) {
                // monopole / single segment: R = com - r, Rp = R + sum ds
                if (MIXED && !(leafnode && d2 < near2)) {
                    float vx, vy, vz;
                    seg_field32((float)ex, (float)ey, (float)ez, (float)(ex + nd.qx),
                                (float)(ey + nd.qy), (float)(ez + nd.qz), f32, vx, vy, vz);
                    ax += vx; ay += vy; az += vz;
                    ++c32;
                } else {
                    double vx, vy, vz;
                    if (seg_field64(ex, ey, ez, ex + nd.qx, ey + nd.qy, ez + nd.qz, f64, vx, vy,
                                    vz)) {
                        ax += vx; ay += vy; az += vz;
                    }
                    ++c64;
                }
                p = nd.skip;
            } else if (nd.count <= leaf) {
                // bucket: every segment of the cell, as count-1 leaves would be
                int s0 = meta[d].start, s1 = s0 + nd.count;
                for (int s = s0; s < s1; ++s) {
                    double bx = sx[s] - qx, by = sy[s] - qy, bz = sz[s] - qz;
                    double b2 = bx * bx + by * by + bz * bz;
                    if (b2 < 1e-16) continue;
                    if (MIXED && !(b2 < near2)) {
                        float vx, vy, vz;
                        seg_field32((float)bx, (float)by, (float)bz, (float)(bx + sqx[s]),
                                    (float)(by + sqy[s]), (float)(bz + sqz[s]), f32, vx, vy, vz);
                        ax += vx; ay += vy; az += vz;
                        ++c32;
                    } else {
                        double vx, vy, vz;
                        if (seg_field64(bx, by, bz, bx + sqx[s], by + sqy[s], bz + sqz[s], f64,
                                        vx, vy, vz)) {
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
        if (valid) {
            if (im < 0) {
                px[k] = ax; py[k] = ay; pz[k] = az;
            } else {
                tx += ax; ty += ay; tz += az;
            }
        }
    }
    if (valid) {
        ix[k] = tx; iy[k] = ty; iz[k] = tz;
    }
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
        if (mixed) go(k_bh_group<true, false>);
        else go(k_bh_group<false, false>);
    } else {
        if (mixed) go(k_bh_group<true, true>);
        else go(k_bh_group<false, true>);
    }
}

void launch_bh(Ctx &c, const double *const x[3]) { launch_bh_targets(c, x, c.n_live, nullptr); }

}  // namespace ovg
