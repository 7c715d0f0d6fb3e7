// The sequential part of the reconnection pass (src/vortex_dynamics.c:
// 186-288) on host mirrors of the positions and links: the greedy visit of
// the candidate pairs in (k, l) order, with the tangent and angle test, the
// approach test on the endpoint velocities, the length test and rewiring
// (do_reconnection, :128-175) and the neighbourhood flags that keep a pass
// from undoing itself.  Host-only; kept apart from the device code so that
// it can be timed and checked on its own.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

#include "host_phys.h"

namespace ovg {
namespace host {

// do_reconnection, src/vortex_dynamics.c:128-175, on the host link mirror.
// Returns 0 (no rewiring) or 1/2 for the two rewiring cases.
inline int do_reconnection(const TangleParams &P, const double *X, int *conn, int k, int l)
{
    int kf = conn[2 * k], kr = conn[2 * k + 1], lf = conn[2 * l], lr = conn[2 * l + 1];
    auto plen = [&](int a, int b) {
        double w[3] = {X[3 * (size_t)b], X[3 * (size_t)b + 1], X[3 * (size_t)b + 2]};
        pwrap(P, X + 3 * (size_t)a, w);
        return dist3(X + 3 * (size_t)a, w);
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

// Greedy pass over the candidate pairs (ck[i], cl[i]) in (k, l) order.
//   X: AoS positions of every slot; conn: (forward, reverse) per slot,
//   rewired in place; n: slots.
//   ends: the distinct candidate endpoints; where[p]: p's index in ends;
//   bsp, bsi: Biot-Savart velocity (primary, images) at each endpoint with
//   the links before the pass, 3 per endpoint.
//   exact: apply the field change of every rewired segment (direct mode);
//   otherwise only of those within 25 rec_dist of the point.
//   flag: n zeros, returned zeroed; mark(slot): called for every slot whose
//   links may have changed.
//   st: the status mirror (status, pin_wall per slot) when the tangle has
//   pinned nodes: stencils and velocities as step_node has them there.
// Returns the number of rewirings.
inline int greedy_pass(const TangleParams &P, const double *X, int *conn, int n,
                       const std::vector<int> &ck, const std::vector<int> &cl,
                       const std::vector<int> &ends, const std::vector<int> &where,
                       const double *bsp, const double *bsi, double rec_angle, double rec_dist,
                       bool exact, std::vector<int> &flag, const std::function<void(int)> &mark,
                       const int *st = nullptr)
{
    const int total = (int)ck.size();
    if ((int)flag.size() < n) flag.assign((size_t)n, 0);
    // The pass reaches each endpoint's links, flag, index and five-point stencil
    // (two links either way) through chains of dependent loads into the
    // position and link mirrors, tens of MB that are not in cache.  Touch
    // them first in three rounds of independent prefetches so that the
    // misses overlap; this changes no result.
    {
        const int *cn = conn;
        auto pf_x = [&](int q) {
            __builtin_prefetch(X + 3 * (size_t)q);
            __builtin_prefetch(X + 3 * (size_t)q + 2);
        };
        for (int e : ends) {
            __builtin_prefetch(cn + 2 * (size_t)e);
            __builtin_prefetch(flag.data() + e);
            __builtin_prefetch(where.data() + e);
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
    // Tangent, normal and uncorrected velocity of every endpoint with the
    // links before the pass, computed in parallel by the same functions as
    // the loop below (so bit-identical to what it would compute).  An entry
    // goes stale when a rewiring changes a link its stencil reads: the
    // stencil of q reads the links of q, of next(q) and of prev(q), so after
    // a rewiring every slot within one link (new links) of the slots whose
    // links changed is stale; the loop then computes it afresh, and always
    // when a velocity needs corrections for rewired segments.
    const int m = (int)ends.size();
    std::vector<double> pre(9 * (size_t)m);
    std::vector<char> stale(m, 0);
    {
        auto fill = [&](int a, int b) {
            for (int i = a; i < b; ++i) {
                const int e = ends[i];
                double *t = &pre[9 * (size_t)i], *nr = t + 3, *v = t + 6;
                tangent_normal(P, X, conn, e, t, nr, st);
                double bs[3], vs[3];
                for (int q = 0; q < 3; ++q) bs[q] = bsp[3 * (size_t)i + q] + bsi[3 * (size_t)i + q];
                compose(P, X, conn, e, t, nr, bs, vs, v, st);
            }
        };
        const int hw = (int)std::thread::hardware_concurrency();
        const int nt = std::max(1, std::min({8, hw > 0 ? hw : 1, m / 256}));
        std::vector<std::thread> pool;
        for (int w = 1; w < nt; ++w)
            pool.emplace_back(fill, (int)((long long)m * w / nt), (int)((long long)m * (w + 1) / nt));
        fill(0, (int)((long long)m / nt));
        for (auto &th : pool) th.join();
    }
    auto stale_at = [&](int q) { return q < 0 || where[q] < 0 || stale[where[q]]; };
    auto make_stale = [&](int q) {
        if (q >= 0 && where[q] >= 0) stale[where[q]] = 1;
    };
    auto get_tangent = [&](int q, double *t, double *nr) {
        if (!stale_at(q)) {
            const double *s0 = &pre[9 * (size_t)where[q]];
            for (int a = 0; a < 3; ++a) {
                t[a] = s0[a];
                nr[a] = s0[3 + a];
            }
        } else {
            tangent_normal(P, X, conn, q, t, nr, st);
        }
    };
    std::vector<int> flagged;
    struct Change {
        int seg, old_f, new_f;
    };
    std::vector<Change> changes;
    const double cutoff = std::cos(rec_angle);
    // A rewiring moves a segment pair of extent ~rec_dist, whose field
    // beyond 25 rec_dist is below 1e-4 of the local one and far below the
    // tree's own error.
    const double rc2 = (25 * rec_dist) * (25 * rec_dist);
    auto near = [&](int a, int b) {
        double d2 = 0;
        for (int q = 0; q < 3; ++q) {
            double d = X[3 * (size_t)a + q] - X[3 * (size_t)b + q];
            if (P.periodic[q]) d -= P.L[q] * std::nearbyint(d / P.L[q]);
            d2 += d * d;
        }
        return d2 < rc2;
    };
    // Without `exact`: the rewired segments indexed on a periodic cell grid
    // of cell side >= the cutoff, so a query visits its 27 neighbour cells;
    // the matches are applied in the order of `changes`, as a full scan with
    // the same predicate would.
    int gn[3] = {0, 0, 0};
    bool use_grid = !exact;
    for (int q = 0; q < 3 && use_grid; ++q) {
        const double h = 25 * rec_dist;
        gn[q] = h > 0 ? (int)std::min(128.0, std::floor(P.L[q] / h)) : 0;
        use_grid = gn[q] >= 3 && P.periodic[q];
    }
    std::vector<int> ghead, gnext, found;
    std::vector<uint64_t> gbits;  // occupied cells, one bit each (fits in L1)
    // cell coordinate along q, wrapped into [0, gn[q]) (a position lies
    // within a box length or so of the box: compares, not a division).  Any
    // consistent assignment works: `near` decides every match.
    double gscale[3] = {0, 0, 0};
    for (int q = 0; q < 3 && use_grid; ++q) gscale[q] = gn[q] / P.L[q];
    auto gcoord = [&](int slot, int q) {
        int v = (int)std::floor((X[3 * (size_t)slot + q] - P.lo[q]) * gscale[q]);
        if (v < 0 || v >= gn[q]) {
            v %= gn[q];
            if (v < 0) v += gn[q];
        }
        return v;
    };
    auto gcell = [&](int slot) {
        return (gcoord(slot, 2) * gn[1] + gcoord(slot, 1)) * gn[0] + gcoord(slot, 0);
    };
    if (use_grid) {
        const size_t cells = (size_t)gn[0] * gn[1] * gn[2];
        ghead.assign(cells, -1);
        gbits.assign(cells / 64 + 1, 0);
    }
    auto index_change = [&](int idx) {
        if (!use_grid) return;
        const int cell = gcell(changes[idx].seg);
        gnext.push_back(ghead[cell]);
        ghead[cell] = idx;
        gbits[cell >> 6] |= 1ull << (cell & 63);
    };
    auto velocity = [&](int p, double *tn, double *nr, double *vels) {
        int e = where[p];
        double bs[3];
        for (int a = 0; a < 3; ++a) bs[a] = bsp[3 * (size_t)e + a] + bsi[3 * (size_t)e + a];
        auto apply = [&](const Change &ch) {
            double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
            // direct mode (exact) adds mirror images unreflected, the tree
            // mode reflects them, as their base velocities do
            segment_all_images(P, X, ch.seg, ch.old_f, p, vo, !exact);
            segment_all_images(P, X, ch.seg, ch.new_f, p, vn, !exact);
            for (int a = 0; a < 3; ++a) bs[a] += vn[a] - vo[a];
        };
        // change of the segments rewired so far
        if (use_grid) {
            found.clear();
            // the three periodic neighbour coordinates along each axis
            int nb[3][3];
            for (int q = 0; q < 3; ++q) {
                const int v = gcoord(p, q);
                nb[q][0] = v == 0 ? gn[q] - 1 : v - 1;
                nb[q][1] = v;
                nb[q][2] = v + 1 == gn[q] ? 0 : v + 1;
            }
            for (int dz = 0; dz < 3; ++dz)
                for (int dy = 0; dy < 3; ++dy) {
                    const int row = (nb[2][dz] * gn[1] + nb[1][dy]) * gn[0];
                    for (int dx = 0; dx < 3; ++dx) {
                        const int cell = row + nb[0][dx];
                        if (!((gbits[cell >> 6] >> (cell & 63)) & 1)) continue;
                        for (int i = ghead[cell]; i >= 0; i = gnext[i])
                            if (near(changes[i].seg, p)) found.push_back(i);
                    }
                }
            std::sort(found.begin(), found.end());
            for (int i : found) apply(changes[i]);
        } else {
            found.clear();
            for (int i = 0; i < (int)changes.size(); ++i)
                if (exact || near(changes[i].seg, p)) found.push_back(i);
            for (int i : found) apply(changes[i]);
        }
        if (found.empty() && !stale_at(p)) {
            const double *s0 = &pre[9 * (size_t)where[p] + 6];
            for (int a = 0; a < 3; ++a) vels[a] = s0[a];
            return;
        }
        double vs[3];
        compose(P, X, conn, p, tn, nr, bs, vs, vels, st);
    };
    int nrec = 0;
    int i = 0;
    while (i < total) {
        int k = ck[i];
        int iend = i;
        while (iend < total && ck[iend] == k) ++iend;
        if (flag[k]) {
            i = iend;
            continue;
        }
        // k's links, tangent and velocity stay fixed until a rewiring, which
        // ends this loop
        double tk[3], nk[3], vk[3];
        get_tangent(k, tk, nk);
        bool have_vk = false;
        for (int j = i; j < iend; ++j) {
            int l = cl[j];
            if (flag[l]) continue;
            double w[3] = {X[3 * (size_t)l], X[3 * (size_t)l + 1], X[3 * (size_t)l + 2]};
            pwrap(P, X + 3 * (size_t)k, w);
            double tl[3], nl[3];
            get_tangent(l, tl, nl);
            if (len3(tk) < 0.5 || len3(tl) < 0.5) continue;
            double calpha = dot3(tk, tl) / len3(tk) / len3(tl);
            if (calpha > cutoff) continue;
            double vl[3];
            if (!have_vk) velocity(k, tk, nk, vk);
            have_vk = true;
            velocity(l, tl, nl, vl);
            double dx[3], dv[3];
            for (int a = 0; a < 3; ++a) dx[a] = X[3 * (size_t)k + a] - w[a];
            for (int a = 0; a < 3; ++a) dv[a] = vk[a] - vl[a];
            if (dot3(dx, dv) > 0) continue;
            int knext = conn[2 * k], kprev = conn[2 * k + 1];
            int lnext = conn[2 * l], lprev = conn[2 * l + 1];
            int kind = do_reconnection(P, X, conn, k, l);
            if (!kind) continue;
            for (int sl : {k, l, knext, kprev, lnext, lprev}) mark(sl);
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
            for (int sl : {k, l, knext, kprev, lnext, lprev}) {
                make_stale(sl);
                if (sl >= 0) {
                    make_stale(conn[2 * sl]);
                    make_stale(conn[2 * sl + 1]);
                }
            }
            ++nrec;
            break;
        }
        i = iend;
    }
    for (int sl : flagged) flag[sl] = 0;
    return nrec;
}

}  // namespace host
}  // namespace ovg
