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
#include "host_phys.h"

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

// Every pair (k, l > k) of the 27 neighbouring cells that passes the
// reference predicate is appended at an atomic counter (order fixed later
// by a sort); beyond cap only the count grows.  The distance test
// dist = sqrt(e.e) > rec_dist is decided on e.e against rec_dist^2 with a
// relative margin of 2^-40 (far wider than the roundings of e.e, rec_dist^2
// and the square root) and evaluated exactly only inside that margin.
__global__ void k_candidates(int n_live, const int *__restrict__ live,
                             const int *__restrict__ sorted, const double *__restrict__ X,
                             const double *__restrict__ Y, const double *__restrict__ Z,
                             const int *__restrict__ fwd, const int *__restrict__ rev,
                             const TangleParams *__restrict__ Pp, Grid g, double rec_dist,
                             const int *__restrict__ start, const int *__restrict__ end,
                             int *__restrict__ count, int2 *__restrict__ out, int cap)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    const TangleParams &P = *Pp;
    const double r2 = rec_dist * rec_dist;
    const double far2 = r2 * (1.0 + 0x1.0p-40), near2 = r2 * (1.0 - 0x1.0p-40);
    int k = live[i];
    double xk = X[k], yk = Y[k], zk = Z[k];
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
                    int l = sorted[s];
                    if (l <= k || l == fk || l == rk) continue;
                    double wx = X[l], wy = Y[l], wz = Z[l];
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
    radix_sort(c, n, 0, 24);
    int *start = c.rc_start.as<int>(), *end = start + ncell;
    hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, ncell,
                       start, INT_MAX);
    hipLaunchKernelGGL(k_fill_int, dim3((ncell + 255) / 256), dim3(256), 0, c.stream, ncell, end,
                       0);
    hipLaunchKernelGGL(k_cell_ranges, dim3(gs), dim3(bs), 0, c.stream, n,
                       T.keys[0].as<uint64_t>(), start, end);
    int total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c.stream, 1, c.rc_count.as<int>(),
                           0);
        hipLaunchKernelGGL(k_candidates, dim3(gs), dim3(bs), 0, c.stream, n, c.live.as<int>(),
                           T.vals[0].as<int>(), X, Y, Z, c.fwd.as<int>(), c.rev.as<int>(),
                           dparams(c), g, c.params.rec_dist, start, end, c.rc_count.as<int>(),
                           c.rc_pairs.as<int2>(), c.rc_cap);
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

}  // namespace ovg

using namespace ovg;

// do_reconnection, src/vortex_dynamics.c:128-175, on the host link mirror.
// Returns 0 (no rewiring) or 1/2 for the two rewiring cases.
static int do_reconnection(const TangleParams &P, const double *X, int *conn, int k, int l)
{
    int kf = conn[2 * k], kr = conn[2 * k + 1], lf = conn[2 * l], lr = conn[2 * l + 1];
    auto plen = [&](int a, int b) {
        double w[3] = {X[3 * (size_t)b], X[3 * (size_t)b + 1], X[3 * (size_t)b + 2]};
        host::pwrap(P, X + 3 * (size_t)a, w);
        return host::dist3(X + 3 * (size_t)a, w);
    };
    double l_kl = plen(k, l), l_kflr = plen(kf, lr), l_kkf = plen(k, kf), l_llr = plen(l, lr);
    double l_krlf = plen(kr, lf), l_kkr = plen(k, kr), l_llf = plen(l, lf);
    double cf1 = l_kl + l_kflr - l_kkf - l_llr;
    double cf2 = l_kl + l_krlf - l_kkr - l_llf;
    if (cf1 > 0 && cf2 > 0) return 0;
    if (cf1 < cf2) {
        conn[2 * k] = l;
        conn[2 * l + 1] = k;
        conn[2 * kf + 1] = lr;
        conn[2 * lr] = kf;
        return 1;
    }
    conn[2 * l] = k;
    conn[2 * k + 1] = l;
    conn[2 * lf + 1] = kr;
    conn[2 * kr] = lf;
    return 2;
}

namespace ovg {

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
            launch_direct(c, xv, m, c.rc_tmp.as<int>(), op, oi, 0);
            c.params = saved;
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
    std::vector<int> &conn = c.h_conn;
    std::vector<int> &flag = c.rc_flag;  // all 0 between passes
    if ((int)flag.size() < n) flag.assign((size_t)n, 0);
    // The pass reaches each endpoint's links, flag and five-point stencil
    // (two links either way) through chains of dependent loads into the
    // position and link mirrors, tens of MB that are not in cache.  Touch
    // them first in three rounds of independent prefetches so that the
    // misses overlap; this changes no result.
    {
        const int *cn = conn.data();
        auto pf_x = [&](int q) {
            __builtin_prefetch(X + 3 * (size_t)q);
            __builtin_prefetch(X + 3 * (size_t)q + 2);
        };
        for (int e : ends) {
            __builtin_prefetch(cn + 2 * (size_t)e);
            __builtin_prefetch(flag.data() + e);
            pf_x(e);
        }
        for (int e : ends)
            for (int q : {cn[2 * (size_t)e], cn[2 * (size_t)e + 1]})
                if (q >= 0) {
                    __builtin_prefetch(cn + 2 * (size_t)q);
                    __builtin_prefetch(flag.data() + q);
                    pf_x(q);
                }
        for (int e : ends) {
            const int f = cn[2 * (size_t)e], r = cn[2 * (size_t)e + 1];
            if (f >= 0 && cn[2 * (size_t)f] >= 0) pf_x(cn[2 * (size_t)f]);
            if (r >= 0 && cn[2 * (size_t)r + 1] >= 0) pf_x(cn[2 * (size_t)r + 1]);
        }
    }
    std::vector<int> flagged;
    struct Change {
        int seg, old_f, new_f;
    };
    std::vector<Change> changes;
    const double cutoff = std::cos(c.params.rec_angle);
    // Direct mode applies the change of every rewired segment (exact); tree
    // mode only those within 25 rec_dist of the point: a rewiring moves a
    // segment pair of extent ~rec_dist, whose field beyond that radius is
    // below 1e-4 of the local one and far below the tree's own error.
    const bool exact = c.params.rec_velocity != OVG_RECVEL_TREE;
    const double rc2 = (25 * c.params.rec_dist) * (25 * c.params.rec_dist);
    auto near = [&](int a, int b) {
        double d2 = 0;
        for (int q = 0; q < 3; ++q) {
            double d = X[3 * (size_t)a + q] - X[3 * (size_t)b + q];
            if (P.periodic[q]) d -= P.L[q] * std::nearbyint(d / P.L[q]);
            d2 += d * d;
        }
        return d2 < rc2;
    };
    // Tree mode: the rewired segments indexed on a periodic cell grid of
    // cell side >= the cutoff, so a query visits its 27 neighbour cells;
    // the matches are applied in the order of `changes`, as a full scan
    // with the same predicate would.
    int gn[3] = {0, 0, 0};
    bool use_grid = !exact;
    for (int q = 0; q < 3 && use_grid; ++q) {
        const double h = 25 * c.params.rec_dist;
        gn[q] = h > 0 ? (int)std::min(128.0, std::floor(P.L[q] / h)) : 0;
        use_grid = gn[q] >= 3 && P.periodic[q];
    }
    std::vector<int> ghead, gnext, gtouched, found;
    auto gcell = [&](int slot) {
        int ci[3];
        for (int q = 0; q < 3; ++q) {
            const double u = (X[3 * (size_t)slot + q] - P.lo[q]) / P.L[q];
            int v = (int)std::floor(u * gn[q]) % gn[q];
            ci[q] = v < 0 ? v + gn[q] : v;
        }
        return (ci[2] * gn[1] + ci[1]) * gn[0] + ci[0];
    };
    if (use_grid) ghead.assign((size_t)gn[0] * gn[1] * gn[2], -1);
    auto index_change = [&](int idx) {
        if (!use_grid) return;
        const int cell = gcell(changes[idx].seg);
        if (ghead[cell] < 0) gtouched.push_back(cell);
        gnext.push_back(ghead[cell]);
        ghead[cell] = idx;
    };
    auto velocity = [&](int p, double *tn, double *nr, double *vels) {
        int e = where[p];
        double bs[3];
        for (int a = 0; a < 3; ++a) bs[a] = bsp[3 * (size_t)e + a] + bsi[3 * (size_t)e + a];
        auto apply = [&](const Change &ch) {
            double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
            host::segment_all_images(P, X, ch.seg, ch.old_f, p, vo);
            host::segment_all_images(P, X, ch.seg, ch.new_f, p, vn);
            for (int a = 0; a < 3; ++a) bs[a] += vn[a] - vo[a];
        };
        // change of the segments rewired so far
        if (use_grid) {
            found.clear();
            const int pc = gcell(p);
            const int px = pc % gn[0], py = (pc / gn[0]) % gn[1], pz = pc / (gn[0] * gn[1]);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int cx = (px + dx + gn[0]) % gn[0], cy = (py + dy + gn[1]) % gn[1],
                                  cz = (pz + dz + gn[2]) % gn[2];
                        for (int i = ghead[(cz * gn[1] + cy) * gn[0] + cx]; i >= 0; i = gnext[i])
                            if (near(changes[i].seg, p)) found.push_back(i);
                    }
            std::sort(found.begin(), found.end());
            for (int i : found) apply(changes[i]);
        } else {
            for (const Change &ch : changes)
                if (exact || near(ch.seg, p)) apply(ch);
        }
        double vs[3];
        host::compose(P, X, conn.data(), p, tn, nr, bs, vs, vels);
    };
    int nrec = 0;
    int i = 0;
    while (i < total) {
        int k = ck[i];
        int iend = i;
        while (iend < total && ck[iend] == k) ++iend;
        if (flag[k]) { i = iend; continue; }
        // k's links, tangent and velocity stay fixed until a rewiring, which
        // ends this loop
        double tk[3], nk[3], vk[3];
        host::tangent_normal(P, X, conn.data(), k, tk, nk);
        bool have_vk = false;
        for (int j = i; j < iend; ++j) {
            int l = cl[j];
            if (flag[l]) continue;
            double w[3] = {X[3 * (size_t)l], X[3 * (size_t)l + 1], X[3 * (size_t)l + 2]};
            host::pwrap(P, X + 3 * (size_t)k, w);
            double tl[3], nl[3];
            host::tangent_normal(P, X, conn.data(), l, tl, nl);
            if (host::len3(tk) < 0.5 || host::len3(tl) < 0.5) continue;
            double calpha = host::dot3(tk, tl) / host::len3(tk) / host::len3(tl);
            if (calpha > cutoff) continue;
            double vl[3];
            if (!have_vk) velocity(k, tk, nk, vk);
            have_vk = true;
            velocity(l, tl, nl, vl);
            double dx[3], dv[3];
            for (int a = 0; a < 3; ++a) dx[a] = X[3 * (size_t)k + a] - w[a];
            for (int a = 0; a < 3; ++a) dv[a] = vk[a] - vl[a];
            if (host::dot3(dx, dv) > 0) continue;
            int knext = conn[2 * k], kprev = conn[2 * k + 1];
            int lnext = conn[2 * l], lprev = conn[2 * l + 1];
            int kind = do_reconnection(P, X, conn.data(), k, l);
            if (!kind) continue;
            for (int sl : {k, l, knext, kprev, lnext, lprev}) mark_dirty(c, sl);
            if (kind == 1) {
                changes.push_back({k, knext, l});
                index_change((int)changes.size() - 1);
                changes.push_back({lprev, l, knext});
                index_change((int)changes.size() - 1);
            } else {
                changes.push_back({l, lnext, k});
                index_change((int)changes.size() - 1);
                changes.push_back({kprev, k, lnext});
                index_change((int)changes.size() - 1);
            }
            for (int sl : {k, conn[2 * k], conn[2 * k + 1], knext, kprev, l, conn[2 * l],
                           conn[2 * l + 1], lnext, lprev}) {
                flag[sl]++;
                flagged.push_back(sl);
            }
            ++nrec;
            break;
        }
        i = iend;
    }
    *n_rec = nrec;
    for (int i = 0; i < m; ++i) where[ends[i]] = -1;
    for (int sl : flagged) flag[sl] = 0;
    if (nrec) OVG_TRY(commit_topology(c));  // the rewired links only
    return OVG_OK;
}

}  // namespace ovg
