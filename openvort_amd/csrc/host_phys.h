// Host-side FP64 formulas for the sequential part of the reconnection pass
// (src/vortex_dynamics.c:224-284): tangents of the two candidate points,
// their velocity composition and the length test.  Same operation order as
// the device kernels; the library is built with -ffp-contract=off.
#pragma once
#include <cmath>

#include "device_common.h"

namespace ovg {
namespace host {

inline double dot3(const double *u, const double *v)
{
    double o = u[0] * v[0];
    o = o + u[1] * v[1];
    o = o + u[2] * v[2];
    return o;
}
inline double len3(const double *u) { return std::sqrt(dot3(u, u)); }
inline double dist3(const double *u, const double *v)
{
    double d[3] = {u[0] - v[0], u[1] - v[1], u[2] - v[2]};
    return len3(d);
}
inline void cross3(double *r, const double *u, const double *v)
{
    double a0 = u[1] * v[2] - u[2] * v[1];
    double a1 = u[2] * v[0] - u[0] * v[2];
    double a2 = u[0] * v[1] - u[1] * v[0];
    r[0] = a0; r[1] = a1; r[2] = a2;
}

// seg_pwrap, src/boundaries.c:54-83
inline void pwrap(const TangleParams &P, const double *r1, double *o)
{
    for (int a = 0; a < 3; ++a) {
        if (!P.periodic[a]) continue;
        int g = 0;
        while ((o[a] - r1[a]) > P.halfL[a] && g++ < 64) o[a] -= P.L[a];
        while ((o[a] - r1[a]) < -P.halfL[a] && g++ < 128) o[a] += P.L[a];
    }
}

// segment_field body with R = r1 - r, Rp = r2 - r (src/tangle.c:540-559)
inline bool seg_field(const double *R, const double *Rp, double f_bs, double *v)
{
    double lR = len3(R), lP = len3(Rp);
    double dot = dot3(R, Rp);
    if (lR < 1e-8 || lP < 1e-8) return false;
    if (dot / lR / lP < 1e-8) return false;
    double denom = lR * lP * (lR * lP + dot);
    double c[3];
    cross3(c, R, Rp);
    double m = f_bs * (lR + lP) / denom;
    for (int q = 0; q < 3; ++q) v[q] = c[q] * m;
    return true;
}

// Direct Biot-Savart of segment m (forward f) at point p (slot pk), all
// images; the primary box skips the two LIA segments (src/tangle.c:658).
// reflect: mirror images reflected as in the Barnes-Hut path
// (src/tangle.c:644-646); the direct path adds them unreflected (:678-683).
inline void segment_all_images(const TangleParams &P, const double *X, int m, int f, int pk,
                               double *acc, bool reflect)
{
    const double *xm = X + 3 * (size_t)m;
    double r2[3] = {X[3 * (size_t)f], X[3 * (size_t)f + 1], X[3 * (size_t)f + 2]};
    pwrap(P, xm, r2);
    const double *r = X + 3 * (size_t)pk;
    for (int im = -1; im < P.n_images; ++im) {
        if (im < 0 && (m == pk || f == pk)) continue;
        double q[3] = {r[0], r[1], r[2]};
        if (im >= 0) image_point(P, im, r[0], r[1], r[2], q[0], q[1], q[2]);
        double R[3] = {xm[0] - q[0], xm[1] - q[1], xm[2] - q[2]};
        double Rp[3] = {r2[0] - q[0], r2[1] - q[1], r2[2] - q[2]};
        double v[3];
        if (seg_field(R, Rp, P.f_bs, v)) {
            if (reflect && im >= 0) image_reflect(P, im, v[0], v[1], v[2]);
            for (int a = 0; a < 3; ++a) acc[a] += v[a];
        }
    }
}

// step_node, src/tangle.c:155-194, as a position: `where` links from k on
// the host mirrors (st: status, pin_wall per slot).  A pinned node at its
// open end walks the other way instead and the result is mirrored through
// its wall (mirror_shift, src/boundaries.c:159-193), innermost mirror
// first; the device's step_pos in the same steps.
inline void step_pos(const TangleParams &P, const double *X, const int *conn, const int *st,
                     int k, int where, double *o)
{
    int walls[4], nw = 0, i = k;
    for (int guard = 0; where != 0 && guard < 8; ++guard) {
        const int s = st[2 * i];
        const int nx = where > 0 ? conn[2 * i] : conn[2 * i + 1];
        if ((s == OVG_PINNED || s == OVG_PINNED_SLIP) && nx < 0) {
            if (nw < 4) walls[nw++] = st[2 * i + 1];
            where = -where;
            continue;
        }
        i = nx;
        where += where > 0 ? -1 : 1;
    }
    for (int a = 0; a < 3; ++a) o[a] = X[3 * (size_t)i + a];
    if (where != 0) o[0] = o[1] = o[2] = std::nan("");
    for (int j = nw - 1; j >= 0; --j) {
        const int w = walls[j], a = w >> 1;
        const double wall = (w & 1) ? P.hi[a] : P.lo[a];
        o[a] = 2 * wall - o[a];
    }
}

// update_tangent_normal, src/tangle.c:423-502; st == nullptr: FREE links
// only (every neighbour a link away), else pinned ends as step_node has them
inline void tangent_normal(const TangleParams &P, const double *X, const int *conn, int k,
                           double *t, double *nrm, const int *st = nullptr)
{
    double q[4][3];
    const double *s1, *s2, *m1, *m2;
    if (st) {
        step_pos(P, X, conn, st, k, 1, q[0]);
        step_pos(P, X, conn, st, k, 2, q[1]);
        step_pos(P, X, conn, st, k, -1, q[2]);
        step_pos(P, X, conn, st, k, -2, q[3]);
        s1 = q[0]; s2 = q[1]; m1 = q[2]; m2 = q[3];
    } else {
        int i1 = conn[2 * k], i2 = conn[2 * i1], im1 = conn[2 * k + 1], im2 = conn[2 * im1 + 1];
        s1 = X + 3 * (size_t)i1; s2 = X + 3 * (size_t)i2;
        m1 = X + 3 * (size_t)im1; m2 = X + 3 * (size_t)im2;
    }
    const double *s0 = X + 3 * (size_t)k;
    double w[4][3];
    const double *ends[4] = {s2, s1, m1, m2};
    for (int j = 0; j < 4; ++j) {
        for (int a = 0; a < 3; ++a) w[j][a] = ends[j][a];
        pwrap(P, s0, w[j]);
    }
    double w12[3] = {s2[0], s2[1], s2[2]}, wm12[3] = {m2[0], m2[1], m2[2]};
    pwrap(P, s1, w12);
    pwrap(P, m1, wm12);
    double ds[4][3];
    for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 3; ++a) ds[j][a] = w[j][a] - s0[a];
    double d1 = dist3(s0, w[1]);
    double d2 = d1 + dist3(s1, w12);
    double dm1 = dist3(s0, w[2]);
    double dm2 = dm1 + dist3(m1, wm12);
    double dsd[4] = {d2 * (d2 - d1) * (dm1 + d2) * (dm2 + d2),
                     d1 * (d2 - d1) * (dm1 + d1) * (dm2 + d1),
                     dm1 * (dm1 + d1) * (dm1 + d2) * (dm2 - dm1),
                     dm2 * (dm2 + d1) * (dm2 + d2) * (dm2 - dm1)};
    double c1[4] = {-d1 * dm1 * dm2, d2 * dm1 * dm2, -d1 * d2 * dm2, d1 * d2 * dm1};
    double c2[4] = {2 * ((dm1 - d1) * dm2 - d1 * dm1), -2 * ((dm1 - d2) * dm2 - d2 * dm1),
                    2 * ((d2 + d1) * dm2 - d1 * d2), -2 * ((d2 + d1) * dm1 - d1 * d2)};
    for (int i = 0; i < 3; ++i) {
        double tt = 0, nn = 0;
        for (int z = 0; z < 4; ++z) {
            tt += c1[z] / dsd[z] * ds[z][i];
            nn += c2[z] / dsd[z] * ds[z][i];
        }
        t[i] = tt;
        nrm[i] = nn;
    }
}

// update_velocity composition (src/tangle.c:600-726) given the Biot-Savart
// sum bs at slot k; st == nullptr: FREE links only, else the neighbours of
// step_node and the pinned cases (:608-615, :718-725).
inline void compose(const TangleParams &P, const double *X, const int *conn, int k,
                    const double *t, const double *nr, const double *bs, double *vs,
                    double *vels, const int *st = nullptr)
{
    const double *p = X + 3 * (size_t)k;
    double qn[3], qp[3];
    const double *next, *prev;
    if (st) {
        step_pos(P, X, conn, st, k, 1, qn);
        step_pos(P, X, conn, st, k, -1, qp);
        next = qn;
        prev = qp;
    } else {
        next = X + 3 * (size_t)conn[2 * k];
        prev = X + 3 * (size_t)conn[2 * k + 1];
    }
    double wf[3] = {next[0], next[1], next[2]}, wr[3] = {p[0], p[1], p[2]};
    pwrap(P, p, wf);
    pwrap(P, prev, wr);
    double l_next = dist3(p, wf), l_prev = dist3(prev, wr);
    double f = P.kappa * std::log(2 * std::sqrt(l_next * l_prev) / P.sqrtE / P.a0) / 4 / M_PI;
    double c[3];
    cross3(c, t, nr);
    for (int a = 0; a < 3; ++a) vs[a] = c[a] * f;
    for (int a = 0; a < 3; ++a) vs[a] = vs[a] + P.evs[a];
    for (int a = 0; a < 3; ++a) vs[a] = vs[a] + bs[a];
    for (int a = 0; a < 3; ++a) vels[a] = vs[a];
    if (P.use_mf) {
        double dv[3], m[3], nn[3];
        for (int a = 0; a < 3; ++a) dv[a] = P.evn[a] - vs[a];
        cross3(m, t, dv);
        for (int a = 0; a < 3; ++a) vels[a] = vels[a] + m[a] * P.alpha;
        cross3(nn, t, m);
        for (int a = 0; a < 3; ++a) vels[a] = vels[a] + nn[a] * -P.alpha_p;
    }
    const int sk = st ? st[2 * k] : OVG_FREE;
    if (sk == OVG_PINNED) {
        // no self-induced motion, the boundary's velocity
        for (int a = 0; a < 3; ++a) {
            vs[a] = 0;
            vels[a] = P.evb[a];
        }
    } else if (sk == OVG_PINNED_SLIP) {
        // the velocity through the wall removed; boundary_normals
        // (src/boundaries.c:11-18) are +-e_axis, inward
        const int w = st[2 * k + 1], ax = w >> 1;
        double n3[3] = {0, 0, 0};
        n3[ax] = (w & 1) ? -1.0 : 1.0;
        const double nv = dot3(n3, vels);
        for (int a = 0; a < 3; ++a) vels[a] = vels[a] - n3[a] * nv;
    }
}

// Inward unit normal of a wall (boundary_normals, src/boundaries.c:11-18).
inline void wall_normal(int w, double *n)
{
    n[0] = n[1] = n[2] = 0;
    n[w >> 1] = (w & 1) ? -1.0 : 1.0;
}

}  // namespace host
}  // namespace ovg
