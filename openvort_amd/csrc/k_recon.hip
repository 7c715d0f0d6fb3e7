// Reconnection (src/vortex_dynamics.c:128-288).
//
// Candidate scan on the GPU: points are binned into a periodic grid whose
// cells are at least rec_dist wide (at most 256 per axis), sorted by cell
// with the radix sort of k_tree.hip, and each point k tests the points l > k
// of its 27 neighbouring cells (9 rows, each one sorted range) with the
// reference predicate: l not a
// neighbour of k and |seg_pwrap(x_k, x_l)| <= rec_dist, in FP64 with the
// reference's operation order, so the candidate list is bit-identical to
// the reference's all-pairs scan (:197-222) before any reconnection.
//
// The greedy pass then runs on the host in (k, l) order, as the reference
// (:224-284): points flagged by an earlier reconnection are skipped (their
// links may have changed; every other point keeps its original links, so
// the precomputed candidate list stays exact), tangents are recomputed with
// the current links, the approach test uses direct-sum velocities, and the
// length test decides the rewiring (do_reconnection, :128-175).  The
// direct sums for all candidate endpoints come from one GPU launch; after
// each rewiring the host adds the exact change of the two segments whose
// forward link changed.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine_internal.h"
#include "host_recon.h"

namespace ovg {

struct Grid {
    int nc[3];
    double cs[3];
};

static Grid make_grid(const ovg_params &p)
{
    Grid g;
    for (int a = 0; a < 3; ++a) {
        double L = p.box_hi[a] - p.box_lo[a];
        int n = (int)std::floor(L / p.rec_dist);
        if (n > 256) n = 256;
        if (n < 3) n = 1;  // too few cells to separate neighbours: one cell
        g.nc[a] = n;
        g.cs[a] = L / n;
    }
    return g;
}

__device__ __forceinline__ int cell_coord(double x, double lo, double cs, int nc)
{
    int c = (int)floor((x - lo) / cs);
    c %= nc;
    if (c < 0) c += nc;
    return c;
}

__global__ void k_cell_keys(int n_live, const int *__restrict__ live, const double *__restrict__ X,
                            const double *__restrict__ Y, const double *__restrict__ Z,
                            const TangleParams *__restrict__ Pp, Grid g,
                            uint64_t *__restrict__ keys, int *__restrict__ vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    const TangleParams &P = *Pp;
    int k = live[i];
    int cx = cell_coord(X[k], P.lo[0], g.cs[0], g.nc[0]);
    int cy = cell_coord(Y[k], P.lo[1], g.cs[1], g.nc[1]);
    int cz = cell_coord(Z[k], P.lo[2], g.cs[2], g.nc[2]);
    keys[i] = ((uint64_t)cz * g.nc[1] + cy) * g.nc[0] + cx;
    vals[i] = k;
}

__global__ void k_fill_int(int n, int *a, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// start[c] / end[c]: the sorted range of cell c (start = INT_MAX, end = 0
// for an empty cell, from the fills before)
__global__ void k_cell_ranges(int n_live, const uint64_t *__restrict__ keys,
                              int *__restrict__ start, int *__restrict__ end)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int c = (int)keys[i];
    if (i == 0 || (int)keys[i - 1] != c) start[c] = i;
    if (i == n_live - 1 || (int)keys[i + 1] != c) end[c] = i + 1;
}

// The points in cell order: position and slot of sorted point s, the slot
// -1 for a point that is not FREE (never a candidate).
__global__ void k_cell_gather(int n_live, const int *__restrict__ sorted,
                              const double *__restrict__ X, const double *__restrict__ Y,
                              const double *__restrict__ Z, const int *__restrict__ status,
                              double4 *__restrict__ sp, int *__restrict__ sslot)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_live) return;
    const int l = sorted[s];
    sp[s] = make_double4(X[l], Y[l], Z[l], 0.0);
    sslot[s] = status[l] == OVG_FREE ? l : -1;
}

// Every pair (k, l > k) of the 27 neighbouring cells that passes the
// reference predicate is appended at an atomic counter (order fixed later
// by a sort); beyond cap only the count grows.  The distance test
// dist = sqrt(e.e) > rec_dist is decided on e.e against rec_dist^2 with a
// relative margin of 2^-40 (far wider than the roundings of e.e, rec_dist^2
// and the square root) and evaluated exactly only inside that margin.
// Targets and neighbours are taken in cell order (sp, sslot: k_cell_gather),
// so a wavefront's neighbourhoods overlap and each range is read
// contiguously.
__global__ void k_candidates(int n_live, const double4 *__restrict__ sp,
                             const int *__restrict__ sslot, const int *__restrict__ fwd,
                             const int *__restrict__ rev, const TangleParams *__restrict__ Pp,
                             Grid g, double rec_dist, const int *__restrict__ start,
                             const int *__restrict__ end, int *__restrict__ count,
                             int2 *__restrict__ out, int cap)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    const TangleParams &P = *Pp;
    const double r2 = rec_dist * rec_dist;
    const double far2 = r2 * (1.0 + 0x1.0p-40), near2 = r2 * (1.0 - 0x1.0p-40);
    const int k = sslot[i];
    if (k < 0) return;  // pinned ends never reconnect (:201-204)
    const double4 pk = sp[i];
    double xk = pk.x, yk = pk.y, zk = pk.z;
    int fk = fwd[k], rk = rev[k];
    int cx = cell_coord(xk, P.lo[0], g.cs[0], g.nc[0]);
    int cy = cell_coord(yk, P.lo[1], g.cs[1], g.nc[1]);
    int cz = cell_coord(zk, P.lo[2], g.cs[2], g.nc[2]);
    int ry0 = g.nc[1] >= 3 ? -1 : 0, ry1 = g.nc[1] >= 3 ? 1 : 0;
    int rz0 = g.nc[2] >= 3 ? -1 : 0, rz1 = g.nc[2] >= 3 ? 1 : 0;
    // cells x-1 .. x+1 of a row are consecutive ids (one range), or two
    // ranges where the row wraps; a grid of fewer than 3 cells is one cell
    const int nx = g.nc[0];
    int xa0, xa1, xb0 = 0, xb1 = -1;  // cell x ranges [xa0, xa1], [xb0, xb1]
    if (nx < 3) {
        xa0 = xa1 = 0;
    } else if (cx == 0) {
        xa0 = 0; xa1 = 1; xb0 = xb1 = nx - 1;
    } else if (cx == nx - 1) {
        xa0 = nx - 2; xa1 = nx - 1; xb0 = xb1 = 0;
    } else {
        xa0 = cx - 1; xa1 = cx + 1;
    }
    for (int dz = rz0; dz <= rz1; ++dz)
        for (int dy = ry0; dy <= ry1; ++dy)
            for (int part = 0; part < 2; ++part) {
                const int x0 = part ? xb0 : xa0, x1 = part ? xb1 : xa1;
                if (x1 < x0) continue;
                int ay = (cy + dy + g.nc[1]) % g.nc[1];
                int az = (cz + dz + g.nc[2]) % g.nc[2];
                const int row = (az * g.nc[1] + ay) * nx;
                // the cells x0..x1 of a row are consecutive in the sorted order
                int s0 = INT_MAX, s1 = 0;
                for (int x = x0; x <= x1; ++x) {
                    s0 = min(s0, start[row + x]);
                    s1 = max(s1, end[row + x]);
                }
                for (int s = s0; s < s1; ++s) {
                    const int l = sslot[s];  // -1 (not FREE) is below every k
                    if (l <= k || l == fk || l == rk) continue;
                    const double4 pl = sp[s];
                    double wx = pl.x, wy = pl.y, wz = pl.z;
                    pwrap(P, xk, yk, zk, wx, wy, wz);
                    double ex = xk - wx, ey = yk - wy, ez = zk - wz;
                    const double d2 = dot3(ex, ey, ez, ex, ey, ez);
                    if (d2 > far2) continue;
                    if (d2 >= near2 && sqrt(d2) > rec_dist) continue;
                    const int pos = atomicAdd(count, 1);
                    if (pos < cap) out[pos] = make_int2(k, l);
                }
            }
}

int candidates_queue(Ctx &c)
{
    const int n = c.n_live;
    OVG_TRY(c.rc_count.ensure(sizeof(int) * 64));
    if (c.rc_cap == 0) c.rc_cap = 1 << 16;
    OVG_TRY(c.rc_pairs.ensure(sizeof(int2) * (size_t)c.rc_cap));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c.stream, 1, c.rc_count.as<int>(), 0);
    if (n == 0) return OVG_OK;
    Grid g = make_grid(c.params);
    int ncell = g.nc[0] * g.nc[1] * g.nc[2];
    TreeBufs &T = c.tree;
    for (int b = 0; b < 2; ++b) {
        OVG_TRY(T.keys[b].ensure(sizeof(uint64_t) * (n + 1)));
        OVG_TRY(T.vals[b].ensure(sizeof(int) * (n + 1)));
    }
    OVG_TRY(c.rc_start.ensure(sizeof(int) * 2 * (size_t)ncell));
    int bs = 256, gs = (n + bs - 1) / bs;
    ++c.tree_id;  // the sort below reuses the tree's key buffers
    const double *X = c.x[0].as<double>(), *Y = c.x[1].as<double>(), *Z = c.x[2].as<double>();
    hipLaunchKernelGGL(k_cell_keys, dim3(gs), dim3(bs), 0, c.stream, n, c.live.as<int>(), X, Y,
                       Z, dparams(c), g, T.keys[0].as<uint64_t>(), T.vals[0].as<int>());
    OVG_TRY(radix_sort(c, n, 0, 24));
    int *start = c.rc_start.as<int>(), *end = start + ncell;
    hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, ncell,
                       start, INT_MAX);
    hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, ncell, end,
                       0);
    hipLaunchKernelGGL(k_cell_ranges, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), start, end);
    OVG_TRY(c.rc_cell.ensure((sizeof(double4) + sizeof(int)) * (size_t)n + 64));
    double4 *sp = c.rc_cell.as<double4>();
    int *sslot = reinterpret_cast<int *>(sp + n);
    hipLaunchKernelGGL(k_cell_gather, dim3(gs), dim3(bs), 0, c.stream, n, T.vals[0].as<int>(), X,
                       Y, Z, c.status.as<int>(), sp, sslot);
    hipLaunchKernelGGL(k_candidates, dim3(gs), dim3(bs), 0, c.stream, n, sp, sslot,
                       c.fwd.as<int>(), c.rev.as<int>(), dparams(c), g, c.params.rec_dist, start,
                       end, c.rc_count.as<int>(), c.rc_pairs.as<int2>(), c.rc_cap);
    return OVG_OK;
}

// Candidate pairs in the reference's scan order (k ascending, then l).
static int candidates_device(Ctx &c, std::vector<int> &ck, std::vector<int> &cl)
{
    int n = c.n_live;
    ck.clear();
    cl.clear();
    if (n == 0) return OVG_OK;
    Grid g = make_grid(c.params);
    int ncell = g.nc[0] * g.nc[1] * g.nc[2];
    TreeBufs &T = c.tree;
    for (int b = 0; b < 2; ++b) {
        OVG_TRY(T.keys[b].ensure(sizeof(uint64_t) * (n + 1)));
        OVG_TRY(T.vals[b].ensure(sizeof(int) * (n + 1)));
    }
    OVG_TRY(c.rc_start.ensure(sizeof(int) * 2 * (size_t)ncell));
    OVG_TRY(c.rc_count.ensure(sizeof(int) * 64));
    if (c.rc_cap == 0) c.rc_cap = 1 << 16;
    OVG_TRY(c.rc_pairs.ensure(sizeof(int2) * (size_t)c.rc_cap));
    int bs = 256, gs = (n + bs - 1) / bs;
    ++c.tree_id;  // the sort below reuses the tree's key buffers
    const double *X = c.x[0].as<double>(), *Y = c.x[1].as<double>(), *Z = c.x[2].as<double>();
    hipLaunchKernelGGL(k_cell_keys, dim3(gs), dim3(bs), 0, c.stream, n, c.live.as<int>(), X, Y,
                       Z, dparams(c), g, T.keys[0].as<uint64_t>(), T.vals[0].as<int>());
    // sort by cell id (< 2^24): four 8-bit passes (an even count)
    OVG_TRY(radix_sort(c, n, 0, 24));
    int *start = c.rc_start.as<int>(), *end = start + ncell;
    hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, ncell,
                       start, INT_MAX);
    hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, ncell, end,
                       0);
    hipLaunchKernelGGL(k_cell_ranges, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), start, end);
    int total = 0;
    OVG_TRY(c.rc_cell.ensure((sizeof(double4) + sizeof(int)) * (size_t)n + 64));
    double4 *sp = c.rc_cell.as<double4>();
    int *sslot = reinterpret_cast<int *>(sp + n);
    hipLaunchKernelGGL(k_cell_gather, dim3(gs), dim3(bs), 0, c.stream, n, T.vals[0].as<int>(), X,
                       Y, Z, c.status.as<int>(), sp, sslot);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c.stream, 1, c.rc_count.as<int>(),
                           0);
        hipLaunchKernelGGL(k_candidates, dim3(gs), dim3(bs), 0, c.stream, n, sp, sslot,
                           c.fwd.as<int>(), c.rev.as<int>(), dparams(c), g, c.params.rec_dist,
                           start, end, c.rc_count.as<int>(), c.rc_pairs.as<int2>(), c.rc_cap);
        OVG_HIP(hipMemcpyAsync(&total, c.rc_count.p, sizeof(int), hipMemcpyDeviceToHost,
                               c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        if (total <= c.rc_cap) break;
        // more candidates than room: grow and search again
        c.rc_cap = total + total / 4;
        OVG_TRY(c.rc_pairs.ensure(sizeof(int2) * (size_t)c.rc_cap));
    }
    std::vector<int2> pr((size_t)total);
    if (total) {
        OVG_HIP(hipMemcpyAsync(pr.data(), c.rc_pairs.p, sizeof(int2) * (size_t)total,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    std::sort(pr.begin(), pr.end(), [](const int2 &a, const int2 &b) {
        return a.x != b.x ? a.x < b.x : a.y < b.y;
    });
    ck.resize((size_t)total);
    cl.resize((size_t)total);
    for (int i = 0; i < total; ++i) {
        ck[i] = pr[i].x;
        cl[i] = pr[i].y;
    }
    return OVG_OK;
}

int recon_candidates(Ctx &c, int max_pairs, int *h_pairs, int *n_pairs)
{
    std::vector<int> ck, cl;
    OVG_TRY(candidates_device(c, ck, cl));
    const int total = (int)ck.size();
    if (n_pairs) *n_pairs = total;
    int m = total < max_pairs ? total : max_pairs;
    if (h_pairs)
        for (int i = 0; i < m; ++i) {
            h_pairs[2 * i] = ck[i];
            h_pairs[2 * i + 1] = cl[i];
        }
    if (total > max_pairs)
        return set_error(OVG_ELIMIT, "%d candidate pairs exceed max_pairs %d", total, max_pairs);
    return OVG_OK;
}

__global__ void k_gather3(int m, const int *__restrict__ idx, const double *__restrict__ a,
                          const double *__restrict__ b, const double *__restrict__ c,
                          double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int k = idx[i];
    out[3 * i] = a[k];
    out[3 * i + 1] = b[k];
    out[3 * i + 2] = c[k];
}

int reconnect_pass(Ctx &c, double t, int *n_rec)
{
    *n_rec = 0;
    int total = 0;
    // the host pass reads positions: their download overlaps the candidate
    // search and the endpoint velocities
    OVG_TRY(host_positions_begin(c));
    std::vector<int> ck, cl;
    OVG_TRY(candidates_device(c, ck, cl));
    total = (int)ck.size();
    if (total == 0) return OVG_OK;
    int n = c.n;
    TangleParams P = host_tparams(c.params, t);
    // direct Biot-Savart at every candidate endpoint with the original links
    std::vector<int> ends;
    ends.reserve(2 * (size_t)total);
    for (int i = 0; i < total; ++i) {
        ends.push_back(ck[i]);
        ends.push_back(cl[i]);
    }
    std::sort(ends.begin(), ends.end());
    ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
    int m = (int)ends.size();
    std::vector<int> &where = c.rc_where;  // all -1 between passes
    if ((int)where.size() < n) where.assign((size_t)n, -1);
    for (int i = 0; i < m; ++i) where[ends[i]] = i;
    std::vector<double> bsp(3 * (size_t)m), bsi(3 * (size_t)m);
    {
        OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)m));
        OVG_HIP(hipMemcpyAsync(c.rc_tmp.p, ends.data(), sizeof(int) * (size_t)m,
                               hipMemcpyHostToDevice, c.stream));
        const double *xv[3] = {c.x[0].as<double>(), c.x[1].as<double>(), c.x[2].as<double>()};
        push_params(c, t);
        double *op[3] = {c.bs_prim[0].as<double>(), c.bs_prim[1].as<double>(),
                         c.bs_prim[2].as<double>()};
        double *oi[3] = {c.bs_img[0].as<double>(), c.bs_img[1].as<double>(),
                         c.bs_img[2].as<double>()};
        if (c.params.rec_velocity == OVG_RECVEL_TREE && c.params.method == OVG_METHOD_BH) {
            // the step's own Barnes-Hut evaluator, traversed for the endpoints only
            OVG_TRY(build_tree(c, xv));
            if (c.params.traverse == OVG_TRAVERSE_LIST)
                OVG_TRY(launch_bh_list_targets(c, xv, m, c.rc_tmp.as<int>()));
            else
                launch_bh_targets(c, xv, m, c.rc_tmp.as<int>());
        } else {
            // update_velocity(tree = NULL): FP64 direct sum at the endpoints
            launch_prep_segments(c, xv);
            ovg_params saved = c.params;
            c.params.precision = OVG_PREC_FP64;
            const int rd = launch_direct(c, xv, m, c.rc_tmp.as<int>(), op, oi, 0);
            c.params = saved;
            OVG_TRY(rd);
        }
        OVG_TRY(c.partial.ensure(sizeof(double) * 3 * (size_t)m));
        hipLaunchKernelGGL(k_gather3, dim3((m + 255) / 256), dim3(256), 0, c.stream, m,
                           c.rc_tmp.as<int>(), op[0], op[1], op[2], c.partial.as<double>());
        OVG_HIP(hipMemcpyAsync(bsp.data(), c.partial.p, sizeof(double) * 3 * (size_t)m,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        hipLaunchKernelGGL(k_gather3, dim3((m + 255) / 256), dim3(256), 0, c.stream, m,
                           c.rc_tmp.as<int>(), oi[0], oi[1], oi[2], c.partial.as<double>());
        OVG_HIP(hipMemcpyAsync(bsi.data(), c.partial.p, sizeof(double) * 3 * (size_t)m,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    // positions (pinned host copy, AoS)
    const double *X = nullptr;
    OVG_TRY(host_positions(c, &X));
    // greedy pass in (k, l) order
    const int nrec = host::greedy_pass(
        P, X, c.h_conn.data(), n, ck, cl, ends, where, bsp.data(), bsi.data(),
        c.params.rec_angle, c.params.rec_dist, c.params.rec_velocity != OVG_RECVEL_TREE, c.rc_flag,
        [&](int sl) { mark_dirty(c, sl); }, c.has_pinned ? c.h_status.data() : nullptr);
    *n_rec = nrec;
    for (int i = 0; i < m; ++i) where[ends[i]] = -1;
    if (nrec) OVG_TRY(commit_topology(c));  // the rewired links only
    return OVG_OK;
}

}  // namespace ovg
